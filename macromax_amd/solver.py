"""Solve the macroscopic Maxwell / Helmholtz equations with the modified Born series on an MI355X.

Mirrors MacroMax ``solver.py``: ``solve()`` (``:25-89``) and ``Solution`` (``:92-1123``) with the same constructor
arguments, properties and iteration semantics. The iteration ``d = chi [G(chi E + s) - E]``, the accept test
``||d|| < ||d_prev||`` and the divergence guard ``alpha <- alpha_r + 1.5 i alpha_i`` follow ``:1049-1098``; one
iteration is a single fused sequence of gfx950 kernels (``BackEndHip.born_step``) with one host read of two FP64
norms. The electric field is double-buffered: the kernels write ``E + d`` into the second buffer, which becomes
the field only when the update is accepted.

Scope: non-magnetic media (``mu`` a scalar, ``xi = zeta = 0``), isotropic or anisotropic (3x3) permittivity,
scalar or vectorial fields on 1-, 2- and 3-dimensional grids. Magnetic and bi-anisotropic media raise
``NotImplementedError``.
"""
from __future__ import annotations

import logging
import math
from collections.abc import Callable

import numpy as np
import scipy.constants as const
import scipy.optimize
import torch

from . import backend
from .bound import Bound, Electric, Magnetic, PeriodicBound
from .grid import Grid

log = logging.getLogger(__name__)

__all__ = ['solve', 'Solution']


def solve(grid, vectorial=None, wavenumber=1.0, angular_frequency=None, vacuum_wavelength=None,
          current_density=None, source_distribution=None, epsilon=None, xi=0.0, zeta=0.0, mu=1.0,
          refractive_index=None, bound=None, initial_field=0.0, dtype=None,
          callback: Callable = lambda s: s.iteration < 1e4 and s.residue > 1e-4, rtol: float = 0,
          maxiter=math.inf):
    """Construct a Solution and iterate it until ``callback`` returns False (``solver.py:25-89``)."""
    if callback is None:
        def callback(_):
            return _.iteration < maxiter and _.residue > rtol
    return Solution(grid=grid, vectorial=vectorial, wavenumber=wavenumber, angular_frequency=angular_frequency,
                    vacuum_wavelength=vacuum_wavelength, current_density=current_density,
                    source_distribution=source_distribution, epsilon=epsilon, xi=xi, zeta=zeta, mu=mu,
                    refractive_index=refractive_index, bound=bound, initial_field=initial_field,
                    dtype=dtype).solve(callback)


class Solution:
    """State of a Born-series solution; iterate it with ``solve()`` or ``for s in solution``."""

    #: magnetic media: record each iteration's kernels in a HIP graph and replay it (False, the default: launch them
    #: one by one; the results are identical). With the fused passes (8 launches per 2D iteration) and the next
    #: iteration launched ahead, direct launches measured faster (512^2: 0.140 against 0.149 ms per step)
    magnetic_graphs = False
    #: magnetic media without graphs: reduce the norms and copy them to the host on a second stream beside the next
    #: iteration (``mmx_magnetic_iterate_split``)
    magnetic_split_norms = True
    #: problems with at least this much work per iteration, field values (voxels x components) x (non-trivial axes
    #: + 1), reduce the norms and copy them to the host on a second stream beside the next iteration
    #: (``mmx_born_iterate_split``): e.g. 1024^2 scalar and 48^3 vectorial grids and larger. Smaller ones, whose
    #: host launches set the pace (25-30 us per iteration), keep both on the launch stream, which takes fewer host
    #: calls (measured on MI355X: 512^2 scalar 29 against 36 us per iteration, 1024^2 43 against 38, 64^3
    #: vectorial 50 against 44, 256^3 vectorial 824 against 806)
    split_norms_min_work = 1 << 20

    def __init__(self, grid, vectorial=None, wavenumber=1.0, angular_frequency=None, vacuum_wavelength=None,
                 current_density=None, source_distribution=None, epsilon=None, xi=0.0, zeta=0.0, mu=1.0,
                 refractive_index=None, bound: Bound | None = None, initial_field=0.0, dtype=None,
                 chi_storage: str = 'auto'):
        """Arguments as MacroMax's ``Solution`` (``solver.py:93-268``), plus ``chi_storage`` (no reference
        counterpart; how the kernels hold an anisotropic chi): 'auto' picks the most compact exact form, 'uniaxial'
        (u I + w n n^T, 28 bytes per voxel in complex64), 'symmetric' (6 entries, 48 bytes) or 'full' (9 entries,
        72 bytes); a form that does not hold the medium raises ValueError."""
        if chi_storage not in ('auto', 'uniaxial', 'symmetric', 'full'):
            raise ValueError(f"chi_storage must be 'auto', 'uniaxial', 'symmetric' or 'full', not {chi_storage!r}")
        self.__chi_storage = chi_storage
        self.__iteration = 0
        if not isinstance(grid, Grid):
            if np.isscalar(grid[0]):
                grid = [grid]
            grid = Grid.from_ranges(*grid)
        self.__grid = grid.immutable

        if angular_frequency is not None:
            self.__wavenumber = angular_frequency / const.c
        elif vacuum_wavelength is not None:
            self.__wavenumber = 2 * const.pi / vacuum_wavelength
        else:
            self.__wavenumber = wavenumber

        self.__previous_update_norm = math.inf
        self.__field_norm = None

        def func2arr(f):
            if callable(f):
                return f(*self.__grid)
            return f if isinstance(f, torch.Tensor) else np.asarray(f)

        # source (solver.py:174-185)
        if source_distribution is None:
            if current_density is None:
                if vectorial is None:
                    vectorial = True
                current_density = np.zeros((1 + 2 * vectorial, *self.grid.shape),
                                           dtype=dtype if dtype is not None else np.complex128)
            current_density = func2arr(current_density)
            source_distribution = current_density * (-1j * self.angular_frequency * const.mu_0)
        else:
            source_distribution = func2arr(source_distribution)

        if vectorial is None:
            vectorial = source_distribution.ndim > self.grid.ndim and source_distribution.shape[0] == 3
        self.__vectorial = bool(vectorial)

        if bound is None:
            bound = PeriodicBound(self.grid)
        self.__bound = bound

        if dtype is None:
            dtype = np.complex64 if (isinstance(source_distribution, torch.Tensor) and
                                     source_distribution.dtype in (torch.complex64, torch.float32)) else \
                (np.complex128 if isinstance(source_distribution, torch.Tensor) else np.asarray(source_distribution).dtype)
        if not np.issubdtype(dtype, np.complexfloating):
            dtype = np.complex64 if dtype in (np.float16, np.float32) else np.complex128

        # the backend works in units of 1/k0 (solver.py:212)
        self.__BE = backend.load(1 + 2 * self.vectorial, self.grid * self.wavenumber, dtype=dtype)
        BE = self.__BE
        # three field buffers: the current field, the output of the iteration in flight and that of the next one,
        # launched before the current one is decided (discarded if the current update is rejected)
        self.__fields = [BE.allocate_array(fill_value=0) for _ in range(3)]
        self.__cur = 0
        self.__pending = None
        self.__pending_version = None  # version of the pending iteration's input buffer when it was launched
        self.__work = BE.allocate_array(fill_value=0)  # first pass of the next iteration (FFT along axis -1)
        self.__work_valid = False
        self.__norms_dev = [torch.zeros(2, dtype=torch.float64, device=BE.device) for _ in range(2)]
        self.__norms_host = [torch.zeros(2, dtype=torch.float64, pin_memory=True) for _ in range(2)]
        # slot k's event: the iteration's norms are in host slot k
        self.__copy_events = [torch.cuda.Event() for _ in range(2)]
        # the second stream (default priority) of the norm reductions and copies, for problems large enough
        self.__copy_stream = None
        values = int(np.prod(self.grid.shape)) * (3 if self.vectorial else 1)
        if values * (sum(1 for n in self.grid.shape if n > 1) + 1) >= self.split_norms_min_work:
            self.__copy_stream = torch.cuda.Stream(device=BE.device)
            for t in self.__norms_dev:
                t.record_stream(self.__copy_stream)  # not reused for other tensors before the copies end
        self.__field_fp64 = None
        self.__refine_cache = None  # (FP64 field, its version, its FP64 residual source, relative residue)
        self.__BE64 = None
        self.__mag_graphs = {}  # (input buffer, output buffer) -> graph of one magnetic iteration
        self.__mag_graph_state = None  # (operator version, plan scratch generation) the graphs were recorded for
        self.__mag_pool = None
        self.__mag_norms = None
        self.__mag_stream = None
        self.E = initial_field

        def media_field(f):  # mu, xi, zeta as complex128 [1|3, 1|3, *grid-broadcastable] device tensors
            f = func2arr(f)
            t = f if isinstance(f, torch.Tensor) else torch.as_tensor(np.asarray(f, dtype=np.complex128))
            t = t.to(device=BE.device, dtype=torch.complex128)
            while t.ndim < self.grid.ndim + 2:
                t = t.unsqueeze(0)
            return t.transpose(0, 1) if (t.shape[0] == 1 and t.shape[1] > 1) else t
        mu = media_field(mu)
        xi = media_field(xi)
        zeta = media_field(zeta)

        if epsilon is None:
            n = func2arr(refractive_index) if refractive_index is not None else 1.0
            n = BE.to_matrix_field(np.asarray(n).astype(dtype))
            epsilon = BE.mul(n, n)
            if BE.is_scalar(n) and bool(torch.any(n.real < 0)):
                # a negative refractive index inverts both the permittivity and the permeability (solver.py:230-234)
                sign = (1 - 2 * (n.real < 0).to(torch.float64)).to(torch.complex128)
                mu = mu * sign
                epsilon = epsilon * sign.to(epsilon.dtype)
        else:
            epsilon = func2arr(epsilon)
            epsilon = BE.astype(epsilon if isinstance(epsilon, torch.Tensor) else np.asarray(epsilon).astype(dtype))

        # boundary (solver.py:242-253)
        if isinstance(self.__bound, Electric):
            chi_b = BE.astype(self.__bound.electric_susceptibility)
            if epsilon.ndim >= 2 + self.grid.ndim and (epsilon.shape[0] > 1 or epsilon.shape[1] > 1):
                chi_b = BE.eye * chi_b
            epsilon = BE.astype(BE.to_matrix_field(epsilon) + chi_b)
        if isinstance(self.__bound, Magnetic):
            chi_m = torch.as_tensor(np.asarray(self.__bound.magnetic_susceptibility, dtype=np.complex128),
                                    device=BE.device)
            mu = mu + torch.eye(3, dtype=torch.complex128, device=BE.device).reshape(
                3, 3, *([1] * self.grid.ndim)) * chi_m

        # magnetic: mu not constant, or any bi-(an)isotropy (solver.py:304-307)
        self.__magnetic = not (BE.allclose(xi) and BE.allclose(zeta) and
                               bool(torch.allclose(mu, mu.reshape(-1)[0].expand_as(mu))))
        self.__mag = None
        self.__source_normalized = None
        if self.__magnetic:
            if not self.vectorial:
                raise ValueError('magnetic and bi-(an)isotropic media need a vectorial calculation')
            self.__mu = complex(1.0)
            self.__prepare_magnetic(source_distribution, epsilon, xi, zeta, mu)
        else:
            self.__mu = complex(mu.reshape(-1)[0])
            self.__prepare_preconditioner(source_distribution, epsilon)
        self.__relative_residue = None

    # ------------------------------------------------------------------ preconditioner (solver.py:270-493)
    def __prepare_preconditioner(self, source_distribution, epsilon):
        BE = self.__BE
        eps = BE.to_matrix_field(epsilon)
        if eps.shape[0] != eps.shape[1] or eps.shape[0] not in (1, 3):
            raise ValueError(f'permittivity must be a scalar or 3x3 tensor field, not {tuple(eps.shape[:2])}')
        if eps.shape[0] == 3 and not self.vectorial:
            raise ValueError('an anisotropic permittivity needs a vectorial calculation')
        eps64 = eps.to(torch.complex128)
        max_allowed_gain = math.sqrt(BE.eps)
        # the set-up reductions run as FP64 kernels over the full grid (solver.py:313-401)
        full = tuple(int(_) for _ in self.grid.shape)
        eps_full = eps64.expand(eps64.shape[0], eps64.shape[1], *full).contiguous()
        plan = BE.plan
        h_min, h_max = plan.setup_gain(eps_full)  # eigenvalues of -i/2 (eps - eps^H)
        # the constant permeability counts too (solver.py:329-336); -i/2 (mu - mu^*) = Im(mu); xi = zeta = 0 here
        mu_h = self.__mu.imag
        gain = h_min < -max_allowed_gain or mu_h < -max_allowed_gain   # has_gain, solver.py:321-327
        loss = -h_max < -max_allowed_gain or -mu_h < -max_allowed_gain  # has_gain(-a): the eigenvalues change sign
        transpose = gain and not loss
        if gain and not transpose:
            log.warning('Convergence not guaranteed: the permittivity or permeability has gain.')

        beta = float(np.real(1.0 / self.__mu))

        def sigma_ee(alpha_r):  # calc_sigmaEE, solver.py:392-401
            return plan.setup_sigma(eps_full, float(alpha_r) * beta) ** 0.5 / abs(beta)

        alpha_real, min_value = scipy.optimize.fmin(lambda a: sigma_ee(np.asarray(a).ravel()[0]), 0.0,
                                                    initial_simplex=[[0.0], [1.0]], disp=False, full_output=True,
                                                    ftol=0.01, xtol=0.01, maxiter=100, maxfun=100)[:2]
        alpha = complex(alpha_real[0] + 1.0j * min_value)
        if transpose:
            alpha = alpha.conjugate()
        del eps_full
        alpha = self.__increase_bias_to_limit_kernel_width(alpha)
        log.info(f'Preconditioner constants: alpha = {alpha.real:0.4g} + {alpha.imag:0.4g}i, beta = {beta:0.4g}')
        self.__alpha = alpha
        self.__beta = beta
        self.__eps = eps  # without any alpha; chi is derived from it
        self.__coef_a = 1.0
        self.__coef_c = 0j
        self.__alpha_base = alpha
        # chi = (i/alpha_i)(eps/beta - alpha I), materialised densely for the kernels (solver.py:480-484)
        full = tuple(int(_) for _ in self.grid.shape)
        chi = eps64 * (1.0j / alpha.imag / beta)
        if chi.shape[0] == 1:
            chi = chi - alpha * 1.0j / alpha.imag
        else:
            chi = chi - torch.eye(3, dtype=torch.complex128, device=BE.device).reshape(3, 3, *([1] * self.grid.ndim)) \
                * (alpha * 1.0j / alpha.imag)
        chi = chi.expand(chi.shape[0], chi.shape[1], *full)
        # reciprocal media (eps^T = eps to 8 rounding units of the element type) are stored symmetrically
        symmetric = False
        uniaxial = None
        storage = self.__chi_storage
        if chi.shape[0] == 3 and storage != 'full':
            asym = torch.amax(torch.abs(eps64 - eps64.transpose(0, 1)))
            symmetric = bool(asym <= 8 * torch.finfo(BE.hardware_dtype).eps * torch.amax(torch.abs(eps64)))
            if symmetric:
                chi = 0.5 * (chi + chi.transpose(0, 1))
                if storage != 'symmetric':
                    uniaxial = _uniaxial_planes(chi, BE.hardware_dtype)
            if storage == 'symmetric' and not symmetric:
                raise ValueError("chi_storage='symmetric' needs a reciprocal permittivity (eps^T = eps)")
            if storage == 'uniaxial' and uniaxial is None:
                raise ValueError("chi_storage='uniaxial' needs eps = a I + b n n^T with a real axis n per voxel")
        if chi.shape[0] == 1:
            self.__chi_kind = 1
            self.__chi = chi.to(BE.hardware_dtype).contiguous()
        elif uniaxial is not None:
            # uniaxial medium: chi = u I + w n n^T, stored as [u, w, (n_x, n_y), n_z] (28 bytes per voxel in c64)
            self.__chi_kind = 4
            self.__chi = uniaxial
        elif symmetric:
            # reciprocal medium: chi^T = chi, stored as the 6 distinct entries {xx, xy, xz, yy, yz, zz}
            self.__chi_kind = 3
            self.__chi = torch.stack([chi[0, 0], chi[0, 1], chi[0, 2], chi[1, 1], chi[1, 2], chi[2, 2]]).to(
                BE.hardware_dtype).contiguous()
        else:
            self.__chi_kind = 2
            self.__chi = chi.to(BE.hardware_dtype).contiguous()
        self.__coef = torch.empty(6, dtype=BE.real_dtype, device=BE.device)
        self.__upload_coef()
        self.source_distribution = source_distribution

    def __prepare_magnetic(self, source_distribution, epsilon, xi, zeta, mu):
        """Preconditioner and operators of a magnetic medium (``solver.py:270-493``), ``magnetic.py``."""
        from .magnetic import MagneticOperator
        BE = self.__BE
        eps = BE.to_matrix_field(epsilon).to(torch.complex128)
        full = tuple(int(_) for _ in self.grid.shape)
        max_allowed_gain = math.sqrt(BE.eps)
        for name, f in (('permittivity', eps), ('xi', xi), ('zeta', zeta), ('permeability', mu)):
            lo, _ = BE.plan.setup_gain(f.expand(f.shape[0], f.shape[1], *full).contiguous())
            if lo < -max_allowed_gain:
                log.warning(f'Convergence not guaranteed: the {name} has gain.')
        self.__mag = MagneticOperator(BE, BE.plan, full, self.grid.step, self.wavenumber, eps, xi, zeta, mu,
                                      self.__increase_bias_to_limit_kernel_width)
        self.__alpha = self.__mag.alpha
        self.__alpha_base = self.__alpha
        self.__beta = self.__mag.beta
        self.__eps = eps
        self.__coef_a = 1.0
        self.__coef_c = 0j
        self.__chi_kind = 0
        self.__chi = None
        self.__coef = None
        self.source_distribution = source_distribution

    def __require_non_magnetic(self, what: str):
        if self.__magnetic:
            raise NotImplementedError(f'{what} is available for non-magnetic media only (the magnetic susceptibility '
                                      f'includes curl operators).')

    def __upload_coef(self):
        self.__invalidate_work()
        a = self.__alpha
        vals = [a.real, a.imag, self.__coef_a, self.__coef_c.real, self.__coef_c.imag, 0.0]
        self.__coef.copy_(torch.tensor(vals, dtype=torch.float64))

    def __update_operators(self, alpha_new: complex):
        """Change alpha after a rejected update (``solver.py:496-529``) via chi_eff = a chi + c I, s_eff = a s."""
        a_old = self.__alpha
        A = a_old.imag / alpha_new.imag
        B = 1.0j * (a_old - alpha_new) / alpha_new.imag
        self.__coef_a = self.__coef_a * A
        self.__coef_c = self.__coef_c * A + B
        self.__alpha = alpha_new
        self.__upload_coef()

    def __increase_bias_to_limit_kernel_width(self, alpha, max_kernel_field_residue=0.001):
        """``solver.py:591-627``."""
        offset = alpha.imag + 0.05
        t = self.__bound.thickness
        t = t[t != 0]
        if t.size > 0:
            d_min = np.amin(t)
            min_kappa = np.log(max_kernel_field_residue) / (-self.wavenumber * d_min)
            n_real = np.maximum(0, float(alpha.real) + min_kappa ** 2) ** 0.5
            min_offset = 2 * n_real * min_kappa
            if offset < min_offset:
                offset = min_offset
        return complex(alpha.real, offset)

    # ------------------------------------------------------------------ properties
    @property
    def grid(self) -> Grid:
        return self.__grid

    @property
    def vectorial(self) -> bool:
        return self.__vectorial

    @property
    def dtype(self):
        return self.__BE.numpy_dtype

    @property
    def wavenumber(self) -> float:
        return self.__wavenumber

    @property
    def angular_frequency(self) -> float:
        return self.__wavenumber * const.c

    @property
    def wavelength(self) -> float:
        return 2.0 * const.pi / self.__wavenumber

    @property
    def magnetic(self) -> bool:
        """True for media with a non-constant permeability or bi-(an)isotropy (``solver.py:304-307``)."""
        return self.__magnetic

    @property
    def bound(self) -> Bound:
        return self.__bound

    @property
    def backend(self):
        return self.__BE

    @property
    def alpha(self) -> complex:
        """Current preconditioner constant alpha (in units of k0^2)."""
        return self.__alpha

    @property
    def beta(self) -> float:
        return self.__beta

    @property
    def chi_kind(self) -> int:
        """Storage of chi on the device: 1 isotropic, 2 full 3x3, 3 complex-symmetric 3x3 (6 entries), 4 uniaxial
        (u I + w n n^T with a real unit axis n)."""
        return self.__chi_kind

    @property
    def chi(self) -> torch.Tensor:
        """The susceptibility chi for the alpha chosen at set-up, as a [1,1,*grid] or [3,3,*grid] tensor."""
        self.__require_non_magnetic('chi')
        if self.__chi_kind == 4:
            return _uniaxial_matrix(self.__chi)
        if self.__chi_kind != 3:
            return self.__chi
        c = self.__chi
        return torch.stack([torch.stack([c[0], c[1], c[2]]), torch.stack([c[1], c[3], c[4]]),
                            torch.stack([c[2], c[4], c[5]])])

    @property
    def chi_effective(self) -> torch.Tensor:
        """chi for the current alpha: a chi + c I (the operator the kernels apply)."""
        self.__require_non_magnetic('chi_effective')
        c = self.chi * self.__coef_a
        if self.__coef_c != 0:
            if c.shape[0] == 1:
                c = c + self.__coef_c
            else:
                c = c + self.__BE.eye * self.__coef_c
        return c

    @property
    def source_normalized(self) -> torch.Tensor:
        """s = (i/(alpha_i beta)) S for the current alpha, [C, 1, *grid]."""
        if self.__magnetic:
            return self.__mag.s
        return self.__source_normalized * self.__coef_a

    @property
    def source_distribution(self) -> np.ndarray:
        s = self.source_normalized[:, 0] * (self.__beta / 1.0j * self.__alpha.imag)
        return self.__BE.asnumpy(s)

    @source_distribution.setter
    def source_distribution(self, new_source):
        """Set S; stores s_base = i S / (alpha_i beta) for the set-up alpha (``solver.py:702-720``).

        The kernels use s_eff = a s_base, which equals i S / (alpha_i beta) for the current alpha because
        a = alpha_i(set-up) / alpha_i(current).
        """
        BE = self.__BE
        if not isinstance(new_source, torch.Tensor):
            new_source = np.asarray(new_source).astype(BE.numpy_dtype)
        S = BE.to_matrix_field(new_source)
        full = (1 + 2 * self.vectorial, 1, *[int(_) for _ in self.grid.shape])
        S = S.to(torch.complex128).expand(*full)
        if self.__magnetic:
            self.__mag.set_source(S)
            self.__previous_update_norm = math.inf
            self.__invalidate_work()
            return
        self.__source_normalized = (S * (1.0j / self.__alpha_base.imag / self.__beta)).to(
            BE.hardware_dtype).contiguous()
        self.__source_box = source_support(self.__source_normalized, self.grid.ndim)
        self.__previous_update_norm = math.inf
        self.__invalidate_work()

    @property
    def j(self) -> np.ndarray:
        return self.source_distribution / (-1.0j * self.angular_frequency * const.mu_0)

    @j.setter
    def j(self, new_j):
        self.source_distribution = np.asarray(new_j) * (-1.0j * self.angular_frequency * const.mu_0)

    @property
    def epsilon(self) -> np.ndarray:
        """Permittivity without the bound (``solver.py:732-753``)."""
        eps = self.__eps
        if isinstance(self.__bound, Electric):
            chi_b = self.__BE.astype(self.__bound.electric_susceptibility)
            if eps.shape[0] > 1:
                chi_b = self.__BE.eye * chi_b
            eps = eps - chi_b
        return self.__BE.asnumpy(eps)

    @epsilon.setter
    def epsilon(self, new_value):
        """Set the permittivity (bound added) and rebuild the preconditioner (``solver.py:755-787``)."""
        BE = self.__BE
        src = self.source_distribution
        eps = BE.to_matrix_field(np.asarray(new_value).astype(BE.numpy_dtype))
        if isinstance(self.__bound, Electric):
            chi_b = BE.astype(self.__bound.electric_susceptibility)
            if eps.shape[0] > 1 or eps.shape[1] > 1:
                chi_b = BE.eye * chi_b
            eps = eps + chi_b
        self.__prepare_preconditioner(src[:, np.newaxis], eps)

    @property
    def refractive_index(self) -> np.ndarray:
        return self.epsilon ** 0.5

    @refractive_index.setter
    def refractive_index(self, new_value):
        self.epsilon = np.asarray(new_value) ** 2

    @property
    def E(self) -> np.ndarray:
        """Electric field in SI units, [C, *grid] (``solver.py:811-821``)."""
        return self.__BE.asnumpy(self.__fields[self.__cur][:, 0] / (self.wavenumber ** 2))

    @E.setter
    def E(self, E):
        BE = self.__BE
        val = E if isinstance(E, torch.Tensor) else np.asarray(E)
        t = BE.astype(val) * (self.wavenumber ** 2)
        if t.ndim == self.grid.ndim + 1:  # [C, *grid] -> [C, 1, *grid]
            t = t.unsqueeze(1)
        BE.assign(t, self.__fields[self.__cur])
        self.__field_fp64 = None
        self.__refine_cache = None
        self.__previous_update_norm = math.inf
        self.__field_norm = None
        self.__invalidate_work()

    @property
    def field_array(self) -> torch.Tensor:
        """The k0^2-scaled field on the device, [C, 1, *grid] (read-only view of the current buffer)."""
        return self.__fields[self.__cur]

    # ------------------------------------------------------------------ derived fields (solver.py:834-995)
    # The backend's derivatives are with respect to k0 x (its grid is in units of 1/k0), as in the reference.
    @property
    def B(self) -> np.ndarray:
        """Magnetic field B = curl E / (i c) in SI units, [3, *grid] (``solver.py:834-845``)."""
        BE = self.__BE
        B = BE.curl(BE.astype(self.E[:, np.newaxis])) / (1.0j * const.c)
        return BE.asnumpy(B)[:, 0]

    @property
    def H(self) -> np.ndarray:
        """Magnetising field H = i (-mu^-1 curl E + i mu^-1 zeta E) / (mu_0 c), [3, *grid] (``solver.py:862-889``)."""
        BE = self.__BE
        if self.__magnetic:
            return BE.asnumpy(self.__mag.H(BE.astype(self.E[:, np.newaxis]), const.mu_0 * const.c))[:, 0]
        mu_H = (-1.0j / (const.mu_0 * const.c)) * BE.curl(BE.astype(self.E[:, np.newaxis]))
        return BE.asnumpy(mu_H * self.__beta)[:, 0]  # beta = 1 / mu for non-magnetic media

    @property
    def D(self) -> np.ndarray:
        """Displacement field D = i (k0 curl H - J) / omega, J = S / (-i omega mu_0), [3, *grid]
        (``solver.py:847-860``)."""
        BE = self.__BE
        D = BE.curl(BE.astype(self.H[:, np.newaxis])) * self.wavenumber
        S = BE.astype(self.source_distribution[:, np.newaxis])
        D = D - (1.0 / (1.0j * self.angular_frequency * const.mu_0)) * S
        D = D * (1j / self.angular_frequency)
        return BE.asnumpy(D)[:, 0]

    @property
    def S(self) -> np.ndarray:
        """Time-averaged Poynting vector Re(E x H*) / 2, [3, *grid] (``solver.py:891-903``)."""
        BE = self.__BE
        E = BE.astype(self.E[:, np.newaxis])
        H = BE.astype(self.H[:, np.newaxis])
        return (0.5 * BE.asnumpy(BE.cross(E, H.conj())).real)[:, 0]

    @property
    def divS(self) -> np.ndarray:
        """Divergence of S with respect to k0 x (the reference's backend units), [*grid] (``solver.py:905-912``)."""
        BE = self.__BE
        return BE.asnumpy(BE.div(BE.astype(self.S[:, np.newaxis]))[0, 0])

    @property
    def energy_density(self) -> np.ndarray:
        """Time-averaged energy density (E.D* + B.H*) / 4, [*grid] (``solver.py:914-930``)."""
        E, B, H, D = self.E, self.B, self.H, self.D
        u = np.sum((E * D.conj()).real, axis=0) + np.sum((B * H.conj()).real, axis=0)
        return u * 0.25

    @property
    def stress_tensor(self) -> np.ndarray:
        """Maxwell stress tensor (eps_0 E E* + mu_0 H H* - I (eps_0 |E|^2 + mu_0 |H|^2) / 2) / 2, real part,
        [3, 3, *grid] (``solver.py:932-953``)."""
        BE = self.__BE
        E = BE.astype(self.E[:, np.newaxis])
        H = BE.astype(self.H[:, np.newaxis])
        E2 = torch.sum(E.abs() ** 2, dim=0)
        H2 = torch.sum(H.abs() ** 2, dim=0)
        sigma = const.epsilon_0 * BE.outer(E, E) + const.mu_0 * BE.outer(H, H)
        sigma = sigma - (0.5 * BE.eye) * (const.epsilon_0 * E2 + const.mu_0 * H2)
        return BE.asnumpy(0.5 * sigma).real

    @property
    def f(self) -> np.ndarray:
        """Electromagnetic force density k0 Re(div sigma) in N/m^3, [3, *grid] (``solver.py:955-970``)."""
        BE = self.__BE
        force = BE.real(BE.div(BE.astype(self.stress_tensor))) * self.wavenumber
        return BE.asnumpy(force)[:, 0]

    @property
    def torque(self) -> np.ndarray:
        """Off-diagonal stress components (sigma_yz, sigma_xz, sigma_xy), [3, *grid] (``solver.py:972-991``)."""
        sigma = self.stress_tensor
        return np.concatenate([sigma[1, 2, np.newaxis], sigma[0, 2, np.newaxis], sigma[0, 1, np.newaxis]])

    @property
    def iteration(self) -> int:
        return self.__iteration

    @iteration.setter
    def iteration(self, it: int = 0):
        self.__iteration = it

    @property
    def previous_update_norm(self) -> float:
        return float(self.__previous_update_norm / (self.wavenumber ** 2))

    @property
    def relative_residue(self) -> float:
        """||d_prev|| / ||E|| (``solver.py:1026-1042``)."""
        if self.__relative_residue is None:
            if self.__previous_update_norm > 0:
                if self.__field_norm is None:
                    self.__field_norm = self.__BE.norm(self.__fields[self.__cur])
                self.__relative_residue = self.__previous_update_norm / self.__field_norm
            else:
                self.__relative_residue = 0.0
        return float(self.__relative_residue)

    @property
    def residue(self) -> float:
        return self.relative_residue

    # ------------------------------------------------------------------ iteration (solver.py:1049-1098)
    def __check_pending(self):
        """An iteration launched ahead is void if its input, the current field, was changed in place since then
        (e.g. through ``field_array``): torch counts in-place writes in the tensor's version; the kernels, which
        write through raw pointers, do not change it."""
        if self.__pending is not None and self.__fields[self.__cur]._version != self.__pending_version:
            self.__invalidate_work()

    def __invalidate_work(self):
        """The work buffer no longer holds the first pass of the current field, and an iteration launched ahead
        is void (the field, the source, the medium or alpha changed)."""
        self.__work_valid = False
        self.__pending = None
        self.__refine_cache = None

    def __launch_iteration(self, src: int, dst: int, k: int):
        """Enqueue one fused iteration from field buffer src into dst, its norms into slot k (copied to pinned host
        memory behind an event). Returns (dst, k, event)."""
        BE = self.__BE
        if self.__copy_stream is None:
            BE.born_iterate(self.__fields[src], self.__fields[dst], self.__work, self.__source_normalized, self.__chi,
                            self.__chi_kind, self.__coef, self.__norms_dev[k], s_box=self.__source_box)
            return dst, k, self.__copy_norms(self.__norms_dev[k], k)
        # the norm reduction and its copy to the host run on the copy stream, behind the iteration's passes
        # (partial slot k)
        BE.born_iterate(self.__fields[src], self.__fields[dst], self.__work, self.__source_normalized, self.__chi,
                        self.__chi_kind, self.__coef, self.__norms_dev[k], s_box=self.__source_box,
                        norm_stream=self.__copy_stream, slot=k, norms_host=self.__norms_host[k])
        ev = self.__copy_events[k]
        ev.record(self.__copy_stream)
        return dst, k, ev

    def __copy_norms(self, norms_dev: torch.Tensor, k: int) -> torch.cuda.Event:
        """Copy an iteration's two norms to pinned host slot k behind the launch stream's work. Returns the event
        that marks the copy's completion; slot k and its event are used again two iterations later, after the host
        has waited for it."""
        self.__norms_host[k].copy_(norms_dev, non_blocking=True)
        ev = self.__copy_events[k]
        ev.record(torch.cuda.current_stream(self.__BE.device))
        return ev

    def step(self) -> bool:
        """One Born iteration; returns True if the update was accepted.

        The next iteration is enqueued before this one's norms are read (assuming acceptance, which holds for
        nearly every iteration), so the device never waits for the host's decision; it is discarded when the update
        is rejected."""
        self.__check_pending()
        if self.__magnetic:
            return self.__magnetic_step()
        self.__iteration += 1
        self.__relative_residue = None
        BE = self.__BE
        cur = self.__cur
        if self.__pending is not None:
            out, k, ev = self.__pending
        else:
            if not self.__work_valid:
                BE.born_prologue(self.__fields[cur], self.__work, self.__source_normalized, self.__chi,
                                 self.__chi_kind, self.__coef, s_box=self.__source_box)
            out, k, ev = self.__launch_iteration(cur, (cur + 1) % 3, 0)
        self.__work_valid = True
        # completes this iteration's successor on speculation: input `out`, output the third buffer
        self.__pending = self.__launch_iteration(out, 3 - cur - out, 1 - k)
        self.__pending_version = self.__fields[out]._version
        ev.synchronize()  # waits for this iteration only
        d2, e2 = self.__norms_host[k].tolist()
        current = math.sqrt(d2)
        with np.errstate(divide='ignore', invalid='ignore'):
            rel = np.float64(current) / np.float64(self.__previous_update_norm)
        if rel < 1:
            self.__cur = out
            self.__field_fp64 = None
            self.__refine_cache = None
            self.__previous_update_norm = current
            self.__field_norm = math.sqrt(e2)
            return True
        log.warning(f'The field update is scaled by {rel:0.3f} >= 1 in iteration {self.__iteration}.')
        alpha_new = complex(self.__alpha.real, 1.50 * self.__alpha.imag)
        log.info(f'Increasing the imaginary part of alpha from {self.__alpha.imag:0.3g} to {alpha_new.imag:0.3g}.')
        self.__update_operators(alpha_new)  # also voids the iteration launched ahead and the work buffer
        return False

    def __magnetic_compute(self, cur: int, nxt: int, norms: torch.Tensor):
        """d = chi (G (chi E + s) - E) for E in field buffer cur; buffer nxt = E + d; norms = (|d|^2, |E + d|^2) in
        FP64 (``solver.py:1070-1082``)."""
        self.__mag.iterate(self.__fields[cur], self.__fields[nxt], norms)

    def __magnetic_record(self, cur: int, nxt: int, k: int):
        """Enqueue one magnetic iteration from field buffer cur into nxt, its two squared norms into device slot k.

        With ``magnetic_graphs``, the iteration's launches are recorded once per (input, output, norm slot) in a HIP
        graph (torch.cuda.CUDAGraph, one memory pool for all of them) and replayed. The graphs hold the operator's
        tensors, alpha and the plan's scratch address, so they are recorded again when the operator changes (new
        alpha or source) or the plan's scratch buffer has been reallocated. The first iteration after such a change
        runs eagerly, which also creates every lazily built constant outside the graph's memory pool."""
        norms = self.__magnetic_norms()[k]
        if not self.magnetic_graphs:
            self.__magnetic_compute(cur, nxt, norms)
            return
        plan = self.__BE.plan
        state = (self.__mag.version, plan.get_param('scratch_gen'))
        if state != self.__mag_graph_state:
            self.__mag_graphs = {}
            self.__mag_pool = None
            self.__mag_graph_state = state
            self.__magnetic_compute(cur, nxt, norms)  # eager
            return
        key = (cur, nxt, k, self.__fields[cur].data_ptr(), self.__fields[nxt].data_ptr())
        g = self.__mag_graphs.get(key)
        if g is None:
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, pool=self.__mag_pool):
                self.__magnetic_compute(cur, nxt, norms)
            if self.__mag_pool is None:
                self.__mag_pool = g.pool()
            if plan.get_param('scratch_gen') != state[1]:  # recording grew the scratch: the graph is stale
                self.__mag_graph_state = None
                self.__magnetic_compute(cur, nxt, norms)
                return
            self.__mag_graphs[key] = g
        g.replay()

    def __magnetic_norms(self) -> list:
        """The two device slots of the magnetic iteration's norms (allocated on first use)."""
        if self.__mag_norms is None:
            self.__mag_norms = [torch.zeros(2, dtype=torch.float64, device=self.__BE.device) for _ in range(2)]
        return self.__mag_norms

    def __magnetic_launch(self, cur: int, nxt: int, k: int):
        """Enqueue one magnetic iteration (cur -> nxt) with its norms copied to pinned host slot k behind an event.
        Returns (nxt, k, event)."""
        if self.magnetic_graphs or not self.magnetic_split_norms:
            self.__magnetic_record(cur, nxt, k)
            return nxt, k, self.__copy_norms(self.__mag_norms[k], k)
        # the norm reduction and its copy to the host on a second stream, behind the iteration (partial slot k)
        norms = self.__magnetic_norms()
        if self.__mag_stream is None:
            self.__mag_stream = torch.cuda.Stream(device=self.__BE.device)
            for t in norms:
                t.record_stream(self.__mag_stream)
        self.__mag.iterate(self.__fields[cur], self.__fields[nxt], norms[k], norm_stream=self.__mag_stream, slot=k,
                           norms_host=self.__norms_host[k])
        ev = self.__copy_events[k]
        ev.record(self.__mag_stream)
        return nxt, k, ev

    def __magnetic_step(self) -> bool:
        """One Born iteration with the magnetic susceptibility operator (``solver.py:1064-1098``). As for other
        media, the next iteration is enqueued (assuming acceptance) before this one's norms are read."""
        self.__iteration += 1
        self.__relative_residue = None
        cur = self.__cur
        if self.__pending is not None:
            out, k, ev = self.__pending
        else:
            out, k, ev = self.__magnetic_launch(cur, (cur + 1) % 3, 0)
        self.__pending = self.__magnetic_launch(out, 3 - cur - out, 1 - k)
        self.__pending_version = self.__fields[out]._version
        ev.synchronize()  # waits for this iteration only
        d2, e2 = self.__norms_host[k].tolist()
        current = math.sqrt(d2)
        with np.errstate(divide='ignore', invalid='ignore'):
            rel = np.float64(current) / np.float64(self.__previous_update_norm)
        if rel < 1:
            self.__cur = out
            self.__field_fp64 = None
            self.__refine_cache = None
            self.__previous_update_norm = current
            self.__field_norm = math.sqrt(e2)
            return True
        log.warning(f'The field update is scaled by {rel:0.3f} >= 1 in iteration {self.__iteration}.')
        alpha_new = complex(self.__alpha.real, 1.50 * self.__alpha.imag)
        log.info(f'Increasing the imaginary part of alpha from {self.__alpha.imag:0.3g} to {alpha_new.imag:0.3g}.')
        self.__mag.set_alpha(alpha_new)
        self.__alpha = alpha_new
        self.__pending = None  # computed with the old alpha
        return False

    # ------------------------------------------------------------------ FP64 refinement
    def __fp64_backend(self):
        if self.__BE64 is None:
            from .backend.hip import BackEndHip
            self.__BE64 = BackEndHip(1 + 2 * self.vectorial, self.grid * self.wavenumber, np.complex128,
                                     device=self.__BE.device)
        return self.__BE64

    def __fp64_chi_apply(self, X: torch.Tensor) -> torch.Tensor:
        """chi_eff X in complex128 for X [C, 1, *grid], from chi as the kernels hold it (uniaxial media as
        u_eff X + w_eff n (n . X), 3x3 media entry by entry), without forming a 3x3 field for a uniaxial chi."""
        a, c = self.__coef_a, self.__coef_c
        if self.__chi_kind == 4:
            planes = self.__chi
            real = torch.float32 if planes.dtype == torch.complex64 else torch.float64
            shape = planes.shape[1:]
            u = planes[0].to(torch.complex128) * a + c
            w = planes[1].to(torch.complex128) * a
            n = (planes[2].real.to(torch.float64), planes[2].imag.to(torch.float64),
                 planes[3].view(real).reshape(-1)[:int(np.prod(shape))].reshape(shape).to(torch.float64))
            ndot = w * (n[0] * X[0] + n[1] * X[1] + n[2] * X[2])
            return torch.stack([u * X[i] + n[i] * ndot for i in range(3)])
        chi = self.chi_effective.to(torch.complex128)
        if chi.shape[0] == 1:
            return chi * X
        return torch.stack([chi[i, 0] * X[0] + chi[i, 1] * X[1] + chi[i, 2] * X[2] for i in range(3)])

    def __fp64_residual_source(self, E: torch.Tensor) -> torch.Tensor:
        """s_c = s + chi E - G^-1 E in FP64, for the current alpha: the Born update of E is chi G s_c and the error
        of E solves the iteration with source s_c."""
        BE64 = self.__fp64_backend()
        s = self.source_normalized.to(torch.complex128)
        chiE = self.__fp64_chi_apply(E)
        a = self.__alpha
        F = BE64.ft(E)
        ginv = (BE64.k2.to(torch.complex128) - a) * (1.0j / a.imag)  # 1/g, g = -i alpha_i / (k^2 - alpha)
        if self.vectorial:
            PL = BE64.longitudinal_projection_ft(F)
            F = ginv * (F - PL) + PL * (-1.0j * a / a.imag)  # 1 / (i alpha_i / alpha) on the longitudinal part
        else:
            F = ginv * F
        return s + chiE - BE64.ift(F)

    def __fp64_update_norm(self, s_c: torch.Tensor) -> float:
        BE64 = self.__fp64_backend()
        a = self.__alpha
        F = BE64.ft(s_c)
        g = (-1.0j * a.imag) / (BE64.k2.to(torch.complex128) - a)
        if self.vectorial:
            PL = BE64.longitudinal_projection_ft(F)
            F = g * (F - PL) + PL * (1.0j * a.imag / a)
        else:
            F = g * F
        return float(torch.linalg.vector_norm(self.__fp64_chi_apply(BE64.ift(F))))

    @property
    def field_array_fp64(self) -> torch.Tensor:
        """The field in double precision: the refined field after ``refine()``, else the field converted."""
        f = self.__field_fp64
        return f if f is not None else self.__fields[self.__cur].to(torch.complex128)

    def refine(self, tol: float = 3e-5, max_iter: int = 10000, steps_per_sync: int = 16) -> dict:
        """One mixed-precision refinement sweep (no reference counterpart; the reference iterates in one precision).

        The Born map is affine, so the error e = E* - E of the current field solves the same iteration with the
        source s_c = s + chi E - G^-1 E (evaluated in FP64 with the double-precision kernels); e is solved with the
        complex64 kernels to relative residue ``tol`` and added in FP64. One sweep multiplies the error by about the
        relative accuracy of the correction solve, so a complex64 solution is brought below the single-precision
        floor of the iteration itself. Returns the FP64 relative residues before and after and the iterations used.
        """
        self.__require_non_magnetic('refine')
        from .batch import BatchSolution
        E = self.field_array_fp64.clone()
        cache = self.__refine_cache  # the residual source of the field the last sweep left, if still current
        if cache is not None and cache[0] is self.__field_fp64 and cache[1] == cache[0]._version:
            s_c, res_before = cache[2], cache[3]
        else:
            s_c = self.__fp64_residual_source(E)
            res_before = self.__fp64_update_norm(s_c) / float(torch.linalg.vector_norm(E))
        self.__refine_cache = None
        # source distribution S_c whose normalised, coefficient-scaled form is s_c
        S_c = s_c * (self.__alpha_base.imag * self.__beta / (1.0j * self.__coef_a))
        bt = BatchSolution.from_sources(self, sources=S_c[None]).solve(tol=tol, max_iter=max_iter,
                                                                      steps_per_sync=steps_per_sync, graph=False)
        E = E + bt.field_arrays()[0].to(torch.complex128)
        self.__field_fp64 = E
        self.__fields[self.__cur].copy_(E.to(self.__fields[self.__cur].dtype))
        self.__invalidate_work()
        self.__field_norm = None
        self.__relative_residue = None
        s_c = self.__fp64_residual_source(E)
        res_after = self.__fp64_update_norm(s_c) / float(torch.linalg.vector_norm(E))
        self.__refine_cache = (E, E._version, s_c, res_after)
        return {'fp64_residue_before': res_before, 'fp64_residue_after': res_after,
                'correction_iterations': int(bt.iteration[0]), 'correction_residue': float(bt.residue[0])}

    # ------------------------------------------------------------------ batched solves (macromax_amd.batch)
    def _batch_operands(self) -> dict:
        """Device operands for a batched solve: chi and s for the set-up alpha, the affine coefficients, the
        current field and the iteration bookkeeping."""
        self.__require_non_magnetic('batched and adjoint solves')
        return {'chi': self.__chi, 'chi_kind': self.__chi_kind, 's': self.__source_normalized,
                's_box': self.__source_box,
                'coef': self.__coef, 'field': self.__fields[self.__cur], 'alpha': self.__alpha, 'alpha_base': self.__alpha_base,
                'beta': self.__beta, 'coef_a': self.__coef_a, 'coef_c': self.__coef_c,
                'prev': self.__previous_update_norm, 'iteration': self.__iteration,
                'ncomp': 1 + 2 * self.vectorial}

    def _set_alpha(self, alpha: complex):
        """Use the preconditioner constant ``alpha`` from now on (chi and s rescaled as after a rejected update)."""
        if complex(alpha) != self.__alpha:
            self.__update_operators(complex(alpha))

    def _batch_source(self, source_distribution) -> torch.Tensor:
        """s for the set-up alpha of this medium and another source distribution S, [C, 1, *grid]."""
        BE = self.__BE
        if not isinstance(source_distribution, torch.Tensor):
            source_distribution = np.asarray(source_distribution).astype(BE.numpy_dtype)
        S = BE.to_matrix_field(source_distribution)
        full = (1 + 2 * self.vectorial, 1, *[int(_) for _ in self.grid.shape])
        S = S.to(torch.complex128).expand(*full)
        return (S * (1.0j / self.__alpha_base.imag / self.__beta)).to(BE.hardware_dtype)

    def _batch_result(self, field: torch.Tensor, iteration: int, prev: float, field_norm: float, alpha: complex,
                      coef_a: float, coef_c: complex):
        """Take over the state a batched solve reached for this problem."""
        self.__fields[self.__cur].copy_(field)
        self.__field_fp64 = None
        self.__refine_cache = None
        self.__iteration = int(iteration)
        self.__previous_update_norm = float(prev)
        self.__field_norm = float(field_norm) if field_norm > 0 else None
        self.__relative_residue = None
        self.__alpha = complex(alpha)
        self.__coef_a = float(coef_a)
        self.__coef_c = complex(coef_c)
        self.__upload_coef()

    def __iter__(self):
        self.iteration = 0
        while True:
            self.step()
            yield self

    def solve(self, callback: Callable = lambda _: _.iteration < 1e4 and _.residue > 1e-4, rtol: float = 0,
              maxiter=math.inf):
        if callback is None:
            def callback(_):
                return _.iteration < maxiter and _.residue > rtol
        for sol in self:
            if not callback(sol):
                break
        return self


def _uniaxial_planes(chi: torch.Tensor, dtype) -> torch.Tensor | None:
    """chi [3, 3, *grid] (complex128, symmetric) as chi = u I + w n n^T with a real unit axis n per voxel, stored as
    [4, *grid] planes [u, w, n_x + i n_y, n_z (in the first reals of the plane)] of ``dtype``; None when chi is not of
    that form to within 8 rounding units of ``dtype`` (relative to its largest entry), about the rounding of storing
    chi itself in ``dtype``.

    With A = chi - tr(chi)/3 I = w (n n^T - I/3), A A^H = |w|^2 (n n^T / 3 + I / 9) is real, a n n^T + b I: b follows
    from its trace and the trace of its square, n from the largest column of A A^H - b I, and
    w = 3/2 (n^T chi n - tr(chi) / 3), u = (tr(chi) - w) / 3."""
    shape = chi.shape[2:]
    eye = torch.eye(3, dtype=chi.dtype, device=chi.device).reshape(3, 3, *([1] * len(shape)))
    tr = chi[0, 0] + chi[1, 1] + chi[2, 2]
    A = chi - eye * (tr / 3)
    P = torch.einsum('ij...,kj...->ik...', A, A.conj()).real
    T1 = P[0, 0] + P[1, 1] + P[2, 2]
    T2 = torch.einsum('ij...,ji...->...', P, P)
    b = (4 * T1 - torch.sqrt(torch.clamp(24 * T2 - 8 * T1 ** 2, min=0))) / 12
    Q = P - torch.eye(3, dtype=P.dtype, device=P.device).reshape(3, 3, *([1] * len(shape))) * b
    diag = torch.stack([Q[0, 0], Q[1, 1], Q[2, 2]])
    j = torch.argmax(diag, dim=0)
    col = torch.gather(Q, 1, j.reshape(1, 1, *shape).expand(3, 1, *shape))[:, 0]  # column j of Q
    cj = torch.gather(diag, 0, j.unsqueeze(0))[0]
    nrm = torch.sqrt(torch.clamp(torch.sum(col ** 2, dim=0), min=0))
    scale = torch.amax(torch.abs(chi))
    flat = cj <= 1e-24 * scale ** 2  # isotropic voxel: any axis
    n = torch.where(flat.unsqueeze(0), torch.tensor([0.0, 0.0, 1.0], dtype=P.dtype, device=P.device).reshape(
        3, *([1] * len(shape))), col / torch.where(flat, torch.ones_like(nrm), nrm).unsqueeze(0))
    nc = n.to(chi.dtype)
    ntchin = torch.einsum('i...,ij...,j...->...', nc, chi, nc)
    w = 1.5 * (ntchin - tr / 3)
    u = (tr - w) / 3
    recon = eye * u + w * nc.unsqueeze(0) * nc.unsqueeze(1)
    if not bool(torch.amax(torch.abs(recon - chi)) <= 8 * torch.finfo(dtype).eps * scale):
        return None
    real = torch.float32 if dtype == torch.complex64 else torch.float64
    out = torch.zeros((4, *shape), dtype=dtype, device=chi.device)
    out[0] = u.to(dtype)
    out[1] = w.to(dtype)
    out[2] = torch.complex(n[0], n[1]).to(dtype)
    out[3].view(real).reshape(-1)[:int(np.prod(shape))] = n[2].reshape(-1).to(real)
    return out


def _uniaxial_matrix(planes: torch.Tensor) -> torch.Tensor:
    """[3, 3, *grid] chi from the uniaxial planes of ``_uniaxial_planes``."""
    shape = planes.shape[1:]
    real = torch.float32 if planes.dtype == torch.complex64 else torch.float64
    nz = planes[3].view(real).reshape(-1)[:int(np.prod(shape))].reshape(shape)
    n = torch.stack([planes[2].real, planes[2].imag, nz]).to(planes.dtype)
    eye = torch.eye(3, dtype=planes.dtype, device=planes.device).reshape(3, 3, *([1] * len(shape)))
    return eye * planes[0] + planes[1] * n.unsqueeze(0) * n.unsqueeze(1)


def source_support(s: torch.Tensor, ndim: int):
    """Bounding box (lo, hi) of the nonzero voxels of a [..., *grid] source tensor (empty box: lo == hi)."""
    nz = s != 0
    while nz.ndim > ndim:
        nz = nz.any(dim=0)
    lo, hi = [], []
    for a in range(ndim):
        other = tuple(d for d in range(ndim) if d != a)
        line = nz.any(dim=other) if other else nz
        idx = torch.nonzero(line).ravel()
        if idx.numel() == 0:
            return [0] * ndim, [0] * ndim
        lo.append(int(idx[0]))
        hi.append(int(idx[-1]) + 1)
    return lo, hi


def _eigvals(m: torch.Tensor) -> torch.Tensor:
    from .backend.hip import hermitian3_eigvals
    return hermitian3_eigvals(m)
