"""
CPU oracle for the modified-Born-series iteration of MacroMax (complex128 NumPy restatement).

TEST INFRASTRUCTURE ONLY. Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module, and only as the checker / the timed CPU baseline. The product package
``macromax_amd`` never imports it.

Every function cites the reference file:line it restates (paths relative to the reference repository,
``python/macromax/...``). The arithmetic follows the reference NumPy backend in complex128:

=====================================  =====================================================================
oracle function                         reference
=====================================  =====================================================================
``k_vectors``                           ``utils/ft/grid.py:426-439`` (Grid.f / Grid.k, ifftshifted fftfreq)
``linear_bound_extinction``             ``bound.py:159-183`` (AbsorbingBound.extinction), ``:196-215``
``bound_susceptibility``                ``bound.py:186-193``
``sigma_ee``                            ``solver.py:392-401`` (calc_DeltaEE_beta2 / calc_sigmaEE)
``gain_eigenvalues`` / ``has_gain``      ``solver.py:321-327`` (eigenvalues of -i/2 (a - a^H))
``find_alpha_beta``                     ``solver.py:329-336, 446-467`` (non-magnetic fmin, conjugate alpha for
                                        gain-only media) + ``:591-627`` (bias)
``chi_base``                            ``solver.py:474-484``
``normalized_source``                   ``solver.py:702-720``
``longitudinal_projection_ft``          ``backend/__init__.py:606-649``
``green_op``                            ``solver.py:564-584`` + ``backend/__init__.py:306-319``
``chi_mul``                             ``backend/__init__.py:415-444`` (einsum 'ij...,jk...->ik...')
``BornOracle.step``                     ``solver.py:1064-1098`` (one __iter__ cycle incl. divergence guard)
``BornOracle.relative_residue``         ``solver.py:1026-1042``
``solve_to``                            ``solver.py:1100-1123`` (Solution.solve with a residue callback)
``transposed_k_vectors``                no reference counterpart: k vectors of the transposed operator
``adjoint_gradient``                    no reference counterpart: reciprocity adjoint of the converged field
                                        (derivation in its docstring), pinned by finite differences and by
                                        autograd through ``oracle/born_torch.py``
``stockham_fft``                        model of the GPU kernels' Stockham pass schedule (no reference
                                        counterpart: the reference calls numpy.fft / torch.fft)
``curl``                                ``backend/__init__.py:472-509`` (curl_ft, i k x F, through the FFT)
``magnetic_setup``                      ``solver.py:350-493`` (mu^-1, chi_EH, chi_HE, eps - xi mu^-1 zeta, the
                                        two-parameter fmin over (alpha_r, sqrt(beta)), chi_EE/EH/HE/HH)
``MagneticBornOracle``                  ``solver.py:496-552, 1064-1098`` (curl-term chi operator, iteration)
=====================================  =====================================================================

Pinning (what checks each function against something other than its own formula) is listed in
``tests/test_oracle.py``. Every function above is pinned there; none is "parity unpinned".
"""
from __future__ import annotations

import numpy as np
import scipy.constants as const
import scipy.optimize

__all__ = [
    'k_vectors', 'linear_bound_extinction', 'bound_susceptibility', 'sigma_ee', 'find_alpha_beta',
    'increase_bias', 'gain_eigenvalues', 'has_gain', 'chi_base', 'normalized_source', 'longitudinal_projection_ft',
    'green_op',
    'chi_mul', 'BornOracle', 'stockham_fft', 'plan_radices', 'mat3_max_eig_hermitian', 'setup_problem',
    'solve_to', 'adjoint_gradient', 'transposed_k_vectors', 'curl', 'magnetic_setup', 'MagneticBornOracle',
]


# ----------------------------------------------------------------------------------------------------------
# Grid and boundary (host-side inputs)
# ----------------------------------------------------------------------------------------------------------
def k_vectors(shape, step):
    """k-space sample vectors, one per axis, in ifftshifted (numpy fftfreq) order.

    Follows ``utils/ft/grid.py:426-439``: ``f = Grid(shape, step=1/extent, origin_at_center=False)`` and
    ``k = 2 pi f``. ``step`` is the (k0-scaled) sample pitch.
    """
    return [2.0 * np.pi * np.fft.fftfreq(int(n), d=float(d)) for n, d in zip(shape, step)]


def linear_bound_extinction(shape, step, thickness, max_extinction_coefficient=0.25):
    """Extinction coefficient kappa of a LinearBound on a plaid grid (``bound.py:159-183, 196-215``).

    ``thickness`` and ``max_extinction_coefficient`` broadcast to ``[ndim, 2]`` (axis, front/back).
    Returns a broadcastable array (singleton dims where the axis does not vary), as the reference does.
    """
    ndim = len(shape)
    thickness = np.broadcast_to(np.asarray(thickness, dtype=float), (ndim, 2))
    kappa_max = np.broadcast_to(np.asarray(max_extinction_coefficient, dtype=float), (ndim, 2))
    kappa = 0.0
    for axis in range(ndim):
        rng = (np.arange(shape[axis]) * float(step[axis])).reshape(
            [-1 if _ == axis else 1 for _ in range(ndim)])
        for back_side in range(2):
            t = thickness[axis, back_side] * np.sign(step[axis])
            if not back_side:
                depth = (rng.ravel()[0] + t) - rng
            else:
                depth = rng - (rng.ravel()[-1] - t)
            depth = depth * np.sign(step[axis])
            in_boundary = depth > 0
            if np.any(in_boundary):
                rel_depth = in_boundary * depth / t
                kappa = np.maximum(kappa, kappa_max[axis, back_side] * rel_depth * in_boundary)
    return kappa


def bound_susceptibility(extinction, background_permittivity=1.0):
    """chi_bound = (n_bg + i kappa)^2 - n_bg^2  (``bound.py:186-193``)."""
    n = np.lib.scimath.sqrt(background_permittivity)
    return (n + 1j * np.asarray(extinction)) ** 2 - background_permittivity


# ----------------------------------------------------------------------------------------------------------
# Preconditioner (non-magnetic)
# ----------------------------------------------------------------------------------------------------------
def mat3_max_eig_hermitian(m):
    """Largest |eigenvalue| per voxel of Hermitian matrix fields ``m`` of shape [3,3,...] or [1,1,...].

    The reference uses a Cardano closed form (``backend/__init__.py:670-850``); for Hermitian input its
    result is the exact eigenvalue set, which ``numpy.linalg.eigvalsh`` gives here.
    """
    if m.shape[0] == 1 and m.shape[1] == 1:
        return np.abs(m[0, 0])
    mm = np.moveaxis(np.moveaxis(m, 0, -1), 0, -1)  # [..., 3, 3]
    w = np.linalg.eigvalsh(mm)
    return np.max(np.abs(w), axis=-1)


def sigma_ee(eps, alpha_real, beta):
    """sigma_EE(alpha_r) = sqrt(max lambda((eps - a b)^H (eps - a b))) / |b|  (``solver.py:392-401``)."""
    ab = alpha_real * beta
    eps = np.asarray(eps, dtype=np.complex128)
    if eps.shape[0] == 1 and eps.shape[1] == 1:
        delta2 = np.abs(eps - ab) ** 2
        return float(np.sqrt(np.max(delta2))) / abs(beta)
    d = eps - ab * np.eye(3).reshape(3, 3, *([1] * (eps.ndim - 2)))
    dh = np.conj(np.swapaxes(d, 0, 1))
    delta2 = np.einsum('ij...,jk...->ik...', dh, d)
    return float(np.max(mat3_max_eig_hermitian(delta2))) ** 0.5 / abs(beta)


def increase_bias(alpha, bound_thickness_nonzero, wavenumber, max_kernel_field_residue=0.001):
    """alpha_i <- max(alpha_i + 0.05, 2 n kappa_min)  (``solver.py:591-627``)."""
    offset = alpha.imag + 0.05
    t = np.asarray(bound_thickness_nonzero, dtype=float)
    t = t[t != 0]
    if t.size > 0:
        d_min = np.amin(t)
        min_kappa = np.log(max_kernel_field_residue) / (-wavenumber * d_min)
        n_real = np.maximum(0, float(alpha.real) + min_kappa ** 2) ** 0.5
        min_offset = 2 * n_real * min_kappa
        if offset < min_offset:
            offset = min_offset
    return alpha.real + 1j * offset


def gain_eigenvalues(a):
    """Eigenvalues of the Hermitian part -i/2 (a - a^H) of a matrix field [n,n,...] (or a scalar), all voxels
    (``solver.py:321-327``: a negative eigenvalue is gain, a positive one loss)."""
    a = np.asarray(a, dtype=np.complex128)
    if a.ndim < 2:
        a = a.reshape(1, 1)
    h = -0.5j * (a - np.conj(np.swapaxes(a, 0, 1)))
    if h.shape[0] == 1:
        return np.real(h[0, 0]).ravel()
    m = np.moveaxis(h.reshape(h.shape[0], h.shape[1], -1), -1, 0)
    return np.linalg.eigvalsh(m).ravel()


def has_gain(a, max_allowed_gain):
    """``solver.py:321-327``."""
    return bool(np.any(gain_eigenvalues(a) < -max_allowed_gain))


def find_alpha_beta(eps, mu=1.0, bound_thickness=(), wavenumber=1.0,
                    max_allowed_gain=np.sqrt(np.finfo(np.float64).eps)):
    """Non-magnetic preconditioner constants (``solver.py:446-470``).

    beta = Re(1/mu); alpha_r = fmin(sigma_EE) with simplex [[0],[1]], xtol=ftol=0.01, maxiter=maxfun=100;
    alpha = alpha_r + i sigma_min, conjugated when the medium (eps and the constant mu) has gain and no loss
    (``solver.py:329-336, 464-465``: gain and loss are eigenvalues of -i/2 (a - a^H) below -max_allowed_gain and
    above +max_allowed_gain, the latter sqrt of the machine epsilon of the calculation's dtype); then the bias.
    """
    beta = float(np.real(1.0 / mu))
    gain = has_gain(eps, max_allowed_gain) or has_gain(mu, max_allowed_gain)
    loss = has_gain(-np.asarray(eps), max_allowed_gain) or has_gain(-np.asarray(mu), max_allowed_gain)
    transpose = gain and not loss
    alpha_real, min_value = scipy.optimize.fmin(
        lambda a: sigma_ee(eps, float(np.asarray(a).ravel()[0]), beta), 0.0,
        initial_simplex=[[0.0], [1.0]], disp=False, full_output=True,
        ftol=0.01, xtol=0.01, maxiter=100, maxfun=100)[:2]
    alpha = alpha_real[0] + 1.0j * min_value
    if transpose:
        alpha = np.conj(alpha)
    alpha = increase_bias(alpha, bound_thickness, wavenumber)
    return complex(alpha), beta


def chi_base(eps, alpha, beta):
    """chi = (i/alpha_i) (eps/beta - alpha I)  (``solver.py:480-484``)."""
    eps = np.asarray(eps, dtype=np.complex128)
    chi = eps * (1.0j / alpha.imag / beta)
    if chi.shape[0] == 1:
        chi = chi - alpha * 1.0j / alpha.imag
    else:
        chi = chi - np.eye(3).reshape(3, 3, *([1] * (chi.ndim - 2))) * (alpha * 1.0j / alpha.imag)
    return chi


def normalized_source(S, alpha, beta):
    """s = (i / (alpha_i beta)) S, shape [C,1,...]  (``solver.py:702-720``)."""
    return np.asarray(S, dtype=np.complex128) * (1.0j / alpha.imag / beta)


# ----------------------------------------------------------------------------------------------------------
# The operators of one iteration
# ----------------------------------------------------------------------------------------------------------
def _kb(k, axis, ndim):
    return k.reshape([-1 if _ == axis else 1 for _ in range(ndim)])


def longitudinal_projection_ft(F, k):
    """Pi_L F = k (k . F) / k^2 with Pi_L(k=0) = I on the first ndim components, 0 on the others.

    Follows ``backend/__init__.py:606-649`` (DC stored at ``:625``, division skipped at flat index 0 ``:639``,
    DC restored ``:645``, components >= ndim zeroed ``:646-647``). ``F`` has shape [C,1,*shape].
    """
    shape = F.shape[2:]
    nd = min(len(shape), len(k))
    kb = [_kb(k[a], a, len(shape)) for a in range(nd)]
    proj = np.zeros(shape, dtype=np.complex128)
    for a in range(nd):
        proj = proj + kb[a] * F[a, 0]
    k2 = sum(_ ** 2 for _ in kb)
    k2 = np.broadcast_to(k2, shape)
    flat = proj.ravel()  # copy-free view of a fresh array
    flat[1:] /= k2.ravel()[1:]
    out = np.zeros((max(F.shape[0], nd), 1, *shape), dtype=np.complex128)
    zero = (0,) * len(shape)
    for a in range(nd):
        out[a, 0] = kb[a] * proj
        out[a, 0][zero] = F[a, 0][zero]
    return out


def green_op(X, k, alpha, vectorial, workers=None):
    """G X = ift(g_op(ft(X))) with g = -i alpha_i / (k^2 - alpha)  (``solver.py:564-584``).

    Vectorial: F <- g (F - Pi_L F) + (i alpha_i / alpha) Pi_L F (``solver.py:570-577``).
    FFT over the spatial axes 2.. (``backend/numpy.py:90-94``), unnormalised forward, 1/N inverse: numpy.fft as the
    reference's NumPy backend without pyFFTW; ``workers`` selects scipy.fft with that many threads instead (a tuned
    CPU variant, as pyFFTW's threads would be, ``backend/numpy.py:46-89``).
    """
    axes = tuple(range(2, X.ndim))
    if workers is not None:
        import scipy.fft
        fftn = lambda a: scipy.fft.fftn(a, axes=axes, workers=workers)
        ifftn = lambda a: scipy.fft.ifftn(a, axes=axes, workers=workers)
    else:
        fftn = lambda a: np.fft.fftn(a, axes=axes)
        ifftn = lambda a: np.fft.ifftn(a, axes=axes)
    F = fftn(X)
    shape = X.shape[2:]
    k2 = sum(_kb(k[a], a, len(shape)) ** 2 for a in range(len(shape)))
    g = -1.0j * alpha.imag / (k2 - alpha)
    if vectorial:
        PL = longitudinal_projection_ft(F, k)
        F = g * (F - PL) + PL * (1.0j * alpha.imag / alpha)
    else:
        F = g * F
    return ifftn(F)


def chi_mul(chi, X):
    """Per-voxel chi X: scalar broadcast if chi is [1,1,...], else 3x3 mat-vec (``backend/__init__.py:415-444``)."""
    if chi.shape[0] == 1 and chi.shape[1] == 1:
        return chi * X
    return np.einsum('ij...,jk...->ik...', chi, X)


class BornOracle:
    """The iteration of ``solver.py:1049-1098`` on complex128 arrays, with the divergence guard.

    State: E (k0^2-scaled field, [C,1,*shape]), chi (base, current alpha), s (normalised source), alpha, beta.
    ``step()`` performs one cycle and returns True if the update was accepted.
    """

    def __init__(self, chi, s, alpha, beta, k, vectorial, E0=None, workers=None):
        self.workers = workers
        self.chi = np.asarray(chi, dtype=np.complex128)
        self.s = np.asarray(s, dtype=np.complex128)
        self.alpha = complex(alpha)
        self.beta = float(beta)
        self.k = [np.asarray(_, dtype=float) for _ in k]
        self.vectorial = bool(vectorial)
        self.E = np.zeros(self.s.shape, dtype=np.complex128) if E0 is None else np.array(E0, dtype=np.complex128)
        self.previous_update_norm = np.inf
        self.iteration = 0
        self.last_update_norm = None
        self.accepted = []

    def update_operators(self, alpha_new):
        """``solver.py:496-529``: rescale s and chi for a new alpha."""
        a_old = self.alpha
        self.s = self.s * (a_old.imag / alpha_new.imag)
        chi = self.chi / (1.0j / a_old.imag)
        if chi.shape[0] == 1:
            chi = chi + (a_old - alpha_new)
        else:
            chi = chi + np.eye(3).reshape(3, 3, *([1] * (chi.ndim - 2))) * (a_old - alpha_new)
        self.chi = chi * (1.0j / alpha_new.imag)
        self.alpha = alpha_new

    def update(self):
        """d = chi [G (chi E + s) - E]  (``solver.py:1070-1075``)."""
        d = chi_mul(self.chi, self.E) + self.s
        d = green_op(d, self.k, self.alpha, self.vectorial, self.workers)
        d = d - self.E
        return chi_mul(self.chi, d)

    def step(self):
        self.iteration += 1
        d = self.update()
        current = float(np.linalg.norm(d.ravel()))
        self.last_update_norm = current
        with np.errstate(divide='ignore', invalid='ignore'):
            rel = current / self.previous_update_norm
        if rel < 1:
            self.E = self.E + d
            self.previous_update_norm = current
            self.accepted.append(True)
            return True
        self.update_operators(self.alpha.real + 1.50j * self.alpha.imag)
        self.accepted.append(False)
        return False

    @property
    def relative_residue(self):
        """``solver.py:1026-1042``: ||d_prev|| / ||E|| (0 if ||d_prev|| == 0)."""
        if self.previous_update_norm > 0:
            return self.previous_update_norm / float(np.linalg.norm(self.E.ravel()))
        return 0.0


def solve_to(chi, s, alpha, beta, k, vectorial, tol=1e-12, max_iter=20000, E0=None):
    """Iterate ``BornOracle`` until ``not (residue > tol and iteration < max_iter)`` (``solver.py:1100-1123``)."""
    o = BornOracle(chi, s, alpha, beta, k, vectorial, E0=E0)
    while True:
        o.step()
        if not (o.relative_residue > tol and o.iteration < max_iter):
            return o


def transposed_k_vectors(k):
    """k vectors of the transposed operator: the Nyquist entry of every even-length axis negated.

    With A = F^-1 D(k) F (F the DFT matrix, symmetric), A^T = F^-1 D(k[-n]) F: the multiplier at the reflected
    index. k[-n] = -k[n] except at the Nyquist index n = N/2 of an even axis, where fftfreq stores -pi/d for both,
    so the dyadic cross terms k_i k_j / k^2 change sign on those planes. Negating the Nyquist entries gives the
    multiplier of A^T (D is even in the whole k vector); the scalar Green's function only sees k^2 and is unchanged.
    """
    out = []
    for kk in k:
        kk = np.array(kk, dtype=float)
        if kk.size % 2 == 0 and kk.size > 1:
            kk[kk.size // 2] = -kk[kk.size // 2]
        out.append(kk)
    return out


def adjoint_gradient(chi, alpha, beta, k, vectorial, E, g, tol=1e-12, max_iter=20000):
    """d loss / d eps for a real loss of the converged (k0^2-scaled) field E, by one adjoint solve.

    No reference counterpart (the reference has no gradient; ``examples/scattering_neural_network.py:206-318``
    differentiates an unrolled scalar model). The converged field solves L E = S / beta with
    L = curl curl (or -laplacian) - eps / beta (``solver.py:566-584`` with ``:480-484``, ``:702-720``). For
    eps^T = eps, L^T is L with the Nyquist entries of k negated (``transposed_k_vectors``). With
    g = dL/dRe(E) + i dL/dIm(E): dL = Re <g, dE> and dE = L^-1 (d eps E / beta), so mu = L^-T conj(g) is a forward
    solve of the transposed operator with S' = beta conj(g) (s' = i conj(g) / alpha_i), and
    dL/dRe(eps) + i dL/dIm(eps) = conj(sum_c mu_c E_c) / beta (isotropic eps) or conj(mu_i E_j) / beta (3x3).
    Returns (gradient, adjoint oracle).
    """
    s_adj = np.conj(np.asarray(g, dtype=np.complex128)) * (1.0j / alpha.imag)
    o = solve_to(chi, s_adj, alpha, beta, transposed_k_vectors(k), vectorial, tol, max_iter)
    mu = o.E
    E = np.asarray(E, dtype=np.complex128)
    if chi.shape[0] == 1:
        grad = np.conj(np.sum(mu * E, axis=0, keepdims=True)) / beta  # [1,1,*shape]
    else:
        grad = np.conj(mu[:, None, 0] * E[None, :, 0]) / beta  # [3,3,*shape]
    return grad, o


# ----------------------------------------------------------------------------------------------------------
# Magnetic and bi-(an)isotropic media
# ----------------------------------------------------------------------------------------------------------
def curl(F, k):
    """curl F = ift(i k x ft(F)) of a [3,1,*shape] field, k in units of k0 (``backend/__init__.py:472-509``)."""
    shape = F.shape[2:]
    axes = tuple(range(2, F.ndim))
    G = np.fft.fftn(F, axes=axes)
    kb = [_kb(k[a], a, len(shape)) if a < len(shape) else 0.0 for a in range(3)]
    out = np.zeros((3, 1, *shape), dtype=np.complex128)
    out[0, 0] = 1j * (kb[1] * G[2, 0] - kb[2] * G[1, 0])
    out[1, 0] = 1j * (kb[2] * G[0, 0] - kb[0] * G[2, 0])
    out[2, 0] = 1j * (kb[0] * G[1, 0] - kb[1] * G[0, 0])
    return np.fft.ifftn(out, axes=axes)


def _matrix_field(a, ndim):
    a = np.asarray(a, dtype=np.complex128)
    while a.ndim < ndim + 2:
        a = a[np.newaxis]
    return a


def _as3(a):
    return a if a.shape[0] == 3 else np.eye(3).reshape(3, 3, *([1] * (a.ndim - 2))) * a


def _mmul(a, b):
    if a.shape[0] == 1 and b.shape[0] == 1:
        return a * b
    return np.einsum('ij...,jk...->ik...', _as3(a), _as3(b))


def _max_eig_herm(h):
    """Largest |eigenvalue| over voxels of a Hermitian [n,n,*shape] field (numpy.linalg.eigvalsh)."""
    if h.shape[0] == 1:
        return float(np.amax(np.abs(h)))
    m = np.moveaxis(h.reshape(3, 3, -1), -1, 0)
    return float(np.amax(np.abs(np.linalg.eigvalsh(m))))


def _adj(a):
    return np.conj(np.swapaxes(a, 0, 1))


def magnetic_setup(shape, pitch, wavenumber, eps, xi, zeta, mu, bound_thickness=(), alpha_beta=None):
    """alpha, beta and (chi_EE, chi_EH, chi_HE, chi_HH) of a magnetic medium (``solver.py:350-493``).

    mu^-1 per voxel; chi_EH,beta = -i xi mu^-1, chi_HE,beta = i mu^-1 zeta; eps' = eps - xi mu^-1 zeta;
    sigma_D = |pi / pitch| / k0; sigma_HH(beta) = sqrt(lambda_max((mu^-1 - beta)^H (mu^-1 - beta))) / beta;
    sigma_EE(alpha_r, beta) = sqrt(lambda_max((eps' - alpha_r beta)^H (eps' - alpha_r beta))) / beta;
    fmin over (alpha_r, sqrt(beta)) of beta (sigma_D^2 sigma_HH + sigma_D (sigma_EH + sigma_HE) / beta + sigma_EE)
    + max(0, beta - 1e-3)^2 from [0, 1] with simplex [[0, 1], [1, 1], [0, 0.9]], xtol = ftol = 0.01,
    maxiter 100, maxfun 200; alpha_i = the bracket at the optimum; then the kernel-width bias. ``alpha_beta`` skips
    the search and builds the operators for the given (alpha, beta).
    """
    nd = len(shape)
    eps, xi, zeta, mu = (_matrix_field(_, nd) for _ in (eps, xi, zeta, mu))
    full = lambda a: np.broadcast_to(a, (a.shape[0], a.shape[1], *shape))
    if mu.shape[0] == 1:
        mu_inv = 1.0 / full(mu)
    else:
        m = np.moveaxis(full(mu).reshape(3, 3, -1), -1, 0)
        mu_inv = np.moveaxis(np.linalg.inv(m), 0, -1).reshape(3, 3, *shape)
    chiEH_beta = -1.0j * _mmul(full(xi), mu_inv)
    chiHE_beta = 1.0j * _mmul(mu_inv, full(zeta))
    eps_e = full(eps) - _mmul(full(xi), -1.0j * chiHE_beta)
    eye = lambda a: np.eye(a.shape[0]).reshape(a.shape[0], a.shape[0], *([1] * nd))
    sigma_d = float(np.linalg.norm(2.0 * np.pi / (2.0 * np.broadcast_to(np.asarray(pitch, float), (nd,))))) / wavenumber
    sigma_he = _max_eig_herm(_mmul(_adj(chiHE_beta), chiHE_beta)) ** 0.5
    sigma_eh = _max_eig_herm(_mmul(_adj(chiEH_beta), chiEH_beta)) ** 0.5

    def sigma_hh(beta):
        d = mu_inv - beta * eye(mu_inv)
        return _max_eig_herm(_mmul(_adj(d), d)) ** 0.5 / abs(beta)

    def sigma_ee_m(alpha_r, beta):
        d = eps_e - alpha_r * beta * eye(eps_e)
        return _max_eig_herm(_mmul(_adj(d), d)) ** 0.5 / abs(beta)

    def total(eta, beta):
        return sigma_d ** 2 * sigma_hh(beta) + sigma_d * (sigma_eh + sigma_he) / abs(beta) + sigma_ee_m(eta, beta)

    def target(vec):
        beta = vec[1] ** 2
        return total(vec[0], beta) * beta + max(0.0, beta - 1e-3) ** 2

    if alpha_beta is None:
        vec = scipy.optimize.fmin(target, [0, 1], initial_simplex=[[0, 1], [1, 1], [0, 0.9]], disp=False,
                                  full_output=False, ftol=0.01, xtol=0.01, maxiter=100, maxfun=200)
        beta = float(vec[1] ** 2)
        alpha = complex(vec[0], total(vec[0], beta))
        alpha = increase_bias(alpha, bound_thickness, wavenumber)
    else:
        alpha, beta = complex(alpha_beta[0]), float(alpha_beta[1])
    f = 1.0j / alpha.imag
    chiEE = eps_e * (f / beta) - eye(eps_e) * (alpha * f)
    chiHH = (eye(mu_inv) - mu_inv / beta) * f
    return alpha, beta, (chiEE, chiEH_beta * (f / beta), chiHE_beta * (f / beta), chiHH), mu_inv, chiHE_beta


class MagneticBornOracle:
    """The iteration of ``solver.py:1064-1098`` with chi E = chi_EE E + chi_HE curl E + curl(chi_EH E + chi_HH curl E)
    (``solver.py:532-552``) and the rescaling of ``solver.py:514-529`` after a rejected update."""

    def __init__(self, ops, s, alpha, beta, k, E0=None):
        self.chiEE, self.chiEH, self.chiHE, self.chiHH = (np.asarray(_, dtype=np.complex128) for _ in ops)
        self.s = np.asarray(s, dtype=np.complex128)
        self.alpha = complex(alpha)
        self.beta = float(beta)
        self.k = [np.asarray(_, dtype=float) for _ in k]
        self.E = np.zeros(self.s.shape, dtype=np.complex128) if E0 is None else np.array(E0, dtype=np.complex128)
        self.previous_update_norm = np.inf
        self.iteration = 0

    def chi_op(self, E):
        ED = curl(E, self.k)
        chiE = chi_mul(self.chiEE, E) + chi_mul(self.chiHE, ED)
        chiH = chi_mul(self.chiEH, E) + chi_mul(self.chiHH, ED)
        return chiE + curl(chiH, self.k)

    def update(self):
        d = green_op(self.chi_op(self.E) + self.s, self.k, self.alpha, True)
        return self.chi_op(d - self.E)

    def update_operators(self, alpha_new):
        a_old = self.alpha
        r = a_old.imag / alpha_new.imag
        self.s = self.s * r
        self.chiEH = self.chiEH * r
        self.chiHE = self.chiHE * r
        self.chiHH = self.chiHH * r
        chi = self.chiEE / (1.0j / a_old.imag)
        chi = chi + (a_old - alpha_new) * np.eye(chi.shape[0]).reshape(chi.shape[0], chi.shape[0],
                                                                         *([1] * (chi.ndim - 2)))
        self.chiEE = chi * (1.0j / alpha_new.imag)
        self.alpha = alpha_new

    def step(self):
        self.iteration += 1
        d = self.update()
        current = float(np.linalg.norm(d.ravel()))
        with np.errstate(divide='ignore', invalid='ignore'):
            rel = current / self.previous_update_norm
        if rel < 1:
            self.E = self.E + d
            self.previous_update_norm = current
            return True
        self.update_operators(self.alpha.real + 1.50j * self.alpha.imag)
        return False

    @property
    def relative_residue(self):
        if self.previous_update_norm > 0:
            return self.previous_update_norm / float(np.linalg.norm(self.E.ravel()))
        return 0.0


def setup_problem(shape, pitch, wavenumber, eps, current_density=None, source_distribution=None,
                  bound_thickness=None, bound_kappa=0.25, vectorial=True, mu=1.0,
                  max_allowed_gain=np.sqrt(np.finfo(np.float64).eps)):
    """Build (chi, s, alpha, beta, k, vectorial) from physical inputs as ``Solution.__init__`` does.

    ``solver.py:174-185`` (S = -i w mu0 J), ``:212`` (grid scaled by k0), ``:242-247`` (eps + bound chi),
    ``:270-493`` (preconditioner), ``:702-720`` (source normalisation).
    ``eps`` is [1,1,*shape] or [3,3,*shape] (broadcastable), ``current_density``/``source_distribution`` [C,*shape].
    """
    ndim = len(shape)
    pitch = np.broadcast_to(np.asarray(pitch, dtype=float), (ndim,))
    if source_distribution is None:
        omega = wavenumber * const.c
        source_distribution = np.asarray(current_density) * (-1j * omega * const.mu_0)
    S = np.asarray(source_distribution, dtype=np.complex128)
    C = 3 if vectorial else 1
    S = np.broadcast_to(S.reshape(C, 1, *shape), (C, 1, *shape))
    eps = np.asarray(eps, dtype=np.complex128)
    while eps.ndim < ndim + 2:
        eps = eps[np.newaxis]
    thickness = np.zeros((ndim, 2))
    if bound_thickness is not None:
        thickness = np.broadcast_to(np.asarray(bound_thickness, dtype=float), (ndim, 2))
        kappa = linear_bound_extinction(shape, pitch, thickness, bound_kappa)
        chi_b = bound_susceptibility(kappa)
        if eps.shape[0] > 1:
            chi_b = np.eye(3).reshape(3, 3, *([1] * ndim)) * chi_b
        eps = eps + chi_b
    alpha, beta = find_alpha_beta(eps, mu, thickness.ravel(), wavenumber, max_allowed_gain)
    chi = chi_base(np.broadcast_to(eps, (eps.shape[0], eps.shape[1], *shape)), alpha, beta)
    s = normalized_source(S, alpha, beta)
    k = k_vectors(shape, pitch * wavenumber)
    return chi, s, alpha, beta, k


# ----------------------------------------------------------------------------------------------------------
# Model of the GPU FFT schedule (checks the kernels' index arithmetic on the CPU)
# ----------------------------------------------------------------------------------------------------------
def plan_radices(n, preferred=(8, 4, 2, 3, 5, 7)):
    """Factorisation used by the GPU plan: greedily the largest supported radices first (8, then 4, ...)."""
    radices = []
    m = int(n)
    for r in preferred:
        while m % r == 0 and m > 1:
            radices.append(r)
            m //= r
    if m != 1:
        raise ValueError(f'FFT length {n} has an unsupported prime factor {m}.')
    return radices


def stockham_fft(x, radices=None, inverse=False):
    """Self-sorting Stockham FFT along the last axis, pass by pass exactly as the HIP kernels index it.

    Pass with radix R after the radices with product p: for butterfly i in [0, N/R), k = i mod p,
    inputs x[i + j N/R] (j < R) times twiddle W_N^(j k N/(pR)), an R-point DFT, outputs written to
    (i - k) R + k + j p. Unnormalised in both directions (the 1/N of the inverse is applied by the caller).
    """
    x = np.array(x, dtype=np.complex128)
    n = x.shape[-1]
    if radices is None:
        radices = plan_radices(n)
    sign = 1.0 if inverse else -1.0
    tw = np.exp(sign * 2j * np.pi * np.arange(n) / n)
    p = 1
    for r in radices:
        y = np.empty_like(x)
        m = n // r
        jj = np.arange(r)
        dft = np.exp(sign * 2j * np.pi * np.outer(jj, jj) / r)
        for i in range(m):
            k = i % p
            v = x[..., i + jj * m] * tw[(jj * k * (n // (p * r))) % n]
            out = (i - k) * r + k
            y[..., out + jj * p] = v @ dft.T
        x = y
        p *= r
    return x
