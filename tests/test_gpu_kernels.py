"""Parity of the HIP kernels (through the C ABI) with the oracle and numpy.fft, on the GPU.

Tolerances (complex64 kernels against complex128 references):
- FFT: relative L2 error <= 2e-6 (single precision, O(log N) rounding);
- one Born iteration (field update d, FP64 norms): relative L2 <= 1e-5 (the per-iteration bar);
- complex128 kernels: relative L2 <= 1e-12.
"""
import numpy as np
import pytest

from oracle import born_oracle as ob

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm((np.asarray(a) - np.asarray(b)).ravel()) / max(np.linalg.norm(np.asarray(b).ravel()), 1e-300))


def _plan(shape, step, dtype, device):
    from macromax_amd.backend.hip import Plan
    return Plan(shape, step, dtype, device)


FFT_SHAPES = [(8,), (1024,), (2048,), (5,), (7,), (105,), (5, 10, 20), (64, 96), (56, 40), (96, 96, 96),
              (128, 128, 128), (1, 64, 1), (16, 1, 48), (3, 5, 7), (256, 256), (24, 2, 40),
              # lengths 3 * 2^k (register kernels with a radix-3 stage) on every axis position
              (192,), (384,), (384, 192), (96, 384, 64), (64, 192, 96),
              # lengths 5 * 2^k (register kernels with a radix-5 stage)
              (320,), (160, 80), (80, 160, 64), (64, 320, 80),
              # 640, 768 and 1024 points (register kernels with up to 128 threads per pencil)
              (1024, 64), (768, 640), (64, 1024), (2048, 32), (16, 2048),
              # prime factors above 7 (Bluestein's algorithm) and axes beyond one LDS tile (four-step split)
              (11,), (101,), (64, 13), (13, 17, 11), (2039,), (8192,), (4097,), (2050, 3), (3, 12288), (24, 11, 40)]


@pytest.mark.parametrize('shape', FFT_SHAPES, ids=lambda s: 'x'.join(map(str, s)))
@pytest.mark.parametrize('dtype', ['c64', 'c128'])
def test_fft_matches_numpy(shape, dtype, device):
    import torch
    print(f'fft {shape} {dtype}', flush=True)
    rng = np.random.default_rng(sum(shape))
    nb = 3
    x = rng.standard_normal((nb, *shape)) + 1j * rng.standard_normal((nb, *shape))
    tdt = torch.complex64 if dtype == 'c64' else torch.complex128
    tol = 2e-6 if dtype == 'c64' else 1e-12
    plan = _plan(shape, [1.0] * len(shape), tdt, device)
    xt = torch.as_tensor(x, dtype=tdt, device=device).contiguous()
    out = torch.empty_like(xt)
    plan.fft(xt, out, -1)
    axes = tuple(range(1, 1 + len(shape)))
    assert rel_l2(out.cpu().numpy(), np.fft.fftn(x, axes=axes)) < tol
    back = torch.empty_like(xt)
    plan.fft(out, back, +1)
    assert rel_l2(back.cpu().numpy(), x) < tol
    # in place inverse
    plan.fft(out, out, +1)
    assert rel_l2(out.cpu().numpy(), x) < tol


RADIX_SHAPES = [(105,), (56, 40), (5, 10, 20), (64, 96), (96, 24, 40), (13, 17, 11), (4097,)]


@pytest.mark.parametrize('shape', RADIX_SHAPES, ids=lambda s: 'x'.join(map(str, s)))
@pytest.mark.parametrize('dtype', ['c64', 'c128'])
@pytest.mark.parametrize('order', [1, 2])
def test_fft_radix_orders_match_numpy(shape, dtype, order, device):
    """The LDS-staged transforms with the other radix orders (MMX_PARAM_RADIX_ORDER; the register kernels off)."""
    import torch
    print(f'fft radix order {order} {shape} {dtype}', flush=True)
    rng = np.random.default_rng(sum(shape) + order)
    x = rng.standard_normal((2, *shape)) + 1j * rng.standard_normal((2, *shape))
    tdt = torch.complex64 if dtype == 'c64' else torch.complex128
    tol = 2e-6 if dtype == 'c64' else 1e-12
    plan = _plan(shape, [1.0] * len(shape), tdt, device)
    plan.set_param('reg_fft', 0)
    plan.set_param('radix_order', order)
    assert plan.get_param('radix_order') == order
    xt = torch.as_tensor(x, dtype=tdt, device=device).contiguous()
    out = torch.empty_like(xt)
    plan.fft(xt, out, -1)
    axes = tuple(range(1, 1 + len(shape)))
    assert rel_l2(out.cpu().numpy(), np.fft.fftn(x, axes=axes)) < tol
    plan.fft(out, out, +1)
    assert rel_l2(out.cpu().numpy(), x) < tol


def _born_case(shape, C, kind, nbatch=1, seed=0, chi_shared=True):
    rng = np.random.default_rng(seed)
    V = int(np.prod(shape))

    def crand(*s):
        return rng.standard_normal(s) + 1j * rng.standard_normal(s)
    E = 0.3 * crand(nbatch, C, 1, *shape)
    s = 0.2 * crand(nbatch, C, 1, *shape)
    nch = 1 if chi_shared else nbatch
    if kind == 1:
        chi = 0.3 * crand(nch, 1, 1, *shape)
    elif kind == 4:  # uniaxial: u I + w n n^T with a real unit axis n
        n = rng.standard_normal((nch, 3, *shape))
        n /= np.linalg.norm(n, axis=1, keepdims=True)
        eye = np.eye(3).reshape(1, 3, 3, *([1] * len(shape)))
        chi = 0.3 * crand(nch, 1, 1, *shape) * eye + 0.3 * crand(nch, 1, 1, *shape) * n[:, :, None] * n[:, None, :]
    else:
        chi = 0.2 * crand(nch, 3, 3, *shape)
        if kind == 3:
            chi = 0.5 * (chi + np.swapaxes(chi, 1, 2))  # complex-symmetric
    alphas = [complex(1.1 + 0.2 * b, 0.35 + 0.05 * b) for b in range(nbatch)]
    a = [1.0 - 0.1 * b for b in range(nbatch)]
    c = [complex(0.05 * b, -0.02 * b) for b in range(nbatch)]
    return E, s, chi, alphas, a, c


def _pack(chi, kind, dtype=np.complex128):
    """Device storage of chi: [1,1,*g], [3,3,*g], packed symmetric [6,*g] or uniaxial [u, w, n_x + i n_y, n_z]
    (n_z in the first reals of the last plane), the latter recovered here with numpy.linalg.eigh."""
    if kind == 4:
        shape = chi.shape[2:]
        m = np.moveaxis(chi.reshape(3, 3, -1), -1, 0)  # [V, 3, 3]
        tr = np.trace(m, axis1=1, axis2=2)
        A = m - tr[:, None, None] / 3 * np.eye(3)
        _, vec = np.linalg.eigh((A @ np.conj(np.swapaxes(A, 1, 2))).real)
        n = vec[:, :, 2]
        w = 1.5 * (np.einsum('vi,vij,vj->v', n, m, n) - tr / 3)
        u = (tr - w) / 3
        out = np.zeros((4, n.shape[0]), dtype=dtype)
        out[0], out[1], out[2] = u, w, n[:, 0] + 1j * n[:, 1]
        out[3].view(np.float32 if dtype == np.complex64 else np.float64)[:n.shape[0]] = n[:, 2]
        return out.reshape(4, *shape)
    if kind != 3:
        return chi.astype(dtype)
    return np.stack([chi[0, 0], chi[0, 1], chi[0, 2], chi[1, 1], chi[1, 2], chi[2, 2]]).astype(dtype)


def _oracle_update(E, s, chi, alpha, a, c, k, vectorial):
    eye = np.eye(chi.shape[0]).reshape(chi.shape[0], chi.shape[1], *([1] * (chi.ndim - 2)))
    chi_eff = a * chi + c * eye
    o = ob.BornOracle(chi_eff, a * s, alpha, 1.0, k, vectorial, E0=E)
    d = o.update()
    return d


BORN_CASES = [
    ((1024,), 1, 1), ((1024,), 3, 1), ((1024,), 3, 2), ((2048,), 3, 2),
    ((256, 256), 1, 1), ((64, 96), 3, 1), ((48, 40), 3, 2), ((1, 64), 3, 2),
    ((32, 32, 32), 1, 1), ((24, 32, 40), 3, 1), ((16, 20, 24), 3, 2), ((5, 10, 20), 3, 2),
    ((14, 21, 8), 3, 1), ((8, 1, 12), 3, 2), ((64, 64, 64), 3, 2),
    ((16, 20, 24), 3, 3), ((64, 48), 3, 3), ((512,), 3, 3),
    # lengths 3 * 2^k (register kernels with a radix-3 stage) with every chi kind
    ((96, 96, 96), 3, 4), ((192, 384), 1, 1), ((384, 96), 3, 3), ((96, 192, 64), 3, 2), ((64, 96, 192), 3, 1),
    ((384,), 3, 2), ((192, 64), 1, 1),
    # lengths 5 * 2^k (register kernels with a radix-5 stage)
    ((80, 80, 80), 3, 4), ((160, 320), 1, 1), ((320, 80), 3, 3), ((80, 160, 64), 3, 2), ((320,), 3, 1),
    # 640, 768 and 1024 points
    ((1024, 256), 3, 3), ((640, 768), 1, 1), ((768, 128), 3, 4), ((64, 1024), 3, 2), ((2048, 64), 3, 3),
    ((64, 2048), 1, 1), ((64, 2048), 3, 4), ((32, 2048), 3, 2), ((2048,), 3, 1),
    # power-of-two axes (register-resident kernels) with every chi kind
    ((64, 128, 64), 3, 1), ((128, 64, 256), 3, 3), ((64, 64, 128), 1, 1), ((128, 256), 3, 3), ((256, 64), 3, 2),
    # uniaxial chi storage, LDS-staged and register kernels
    ((16, 20, 24), 3, 4), ((64, 48), 3, 4), ((128, 64, 256), 3, 4), ((128, 256), 3, 4),
    # unfused iteration: prime factors above 7, four-step axes, fused tiles beyond LDS
    ((13, 17, 11), 3, 2), ((101, 64), 1, 1), ((11, 24, 19), 3, 4), ((8192,), 1, 1), ((4097,), 3, 1),
    ((4096,), 3, 1), ((2050, 6), 3, 3),
]


@pytest.mark.parametrize('shape,C,kind', BORN_CASES, ids=lambda v: str(v))
@pytest.mark.parametrize('dtype', ['c64', 'c128'])
def test_born_step_matches_oracle(shape, C, kind, dtype, device):
    import torch
    if dtype == 'c128' and np.prod(shape) > 40000:
        pytest.skip('complex128 covered on the smaller grids')
    print(f'born {shape} C={C} chi_kind={kind} {dtype}', flush=True)
    E, s, chi, alphas, a, c = _born_case(shape, C, kind, seed=len(shape) * 7 + C)
    step = [0.37 + 0.05 * i for i in range(len(shape))]
    k = ob.k_vectors(shape, step)
    tdt = torch.complex64 if dtype == 'c64' else torch.complex128
    rdt = torch.float32 if dtype == 'c64' else torch.float64
    tol = 1e-5 if dtype == 'c64' else 1e-11
    plan = _plan(shape, step, tdt, device)
    Et = torch.as_tensor(E[0], dtype=tdt, device=device).contiguous()
    Eo = torch.zeros_like(Et)
    st = torch.as_tensor(s[0], dtype=tdt, device=device).contiguous()
    ct = torch.as_tensor(_pack(chi[0], kind, np.complex64 if tdt == torch.complex64 else np.complex128),
                         device=device).contiguous()
    coef = torch.tensor([alphas[0].real, alphas[0].imag, a[0], c[0].real, c[0].imag, 0.0], dtype=rdt, device=device)
    norms = torch.zeros(2, dtype=torch.float64, device=device)
    plan.born_step(Et, Eo, st, ct, C, kind, coef, 1, 0, 0, norms)
    torch.cuda.synchronize()
    d_ref = _oracle_update(E[0], s[0], chi[0], alphas[0], a[0], c[0], k, C == 3)
    d_gpu = Eo.cpu().numpy() - E[0]
    assert rel_l2(d_gpu, d_ref) < tol
    dd, ee = norms.cpu().numpy()
    assert abs(dd - np.sum(np.abs(d_ref) ** 2)) / np.sum(np.abs(d_ref) ** 2) < 2 * tol
    en = E[0] + d_ref
    assert abs(ee - np.sum(np.abs(en) ** 2)) / np.sum(np.abs(en) ** 2) < 2 * tol
    # the input field is untouched
    assert np.array_equal(Et.cpu().numpy(), E[0].astype(Et.cpu().numpy().dtype))


@pytest.mark.parametrize('shape,C,kind,dtype', [((24, 40, 56), 3, 4, 'c64'), ((96, 80), 1, 1, 'c64'),
                                                ((16, 20, 24), 3, 2, 'c128'), ((13, 17, 11), 3, 3, 'c64')],
                         ids=lambda v: str(v))
@pytest.mark.parametrize('order', [0, 1, 2])
def test_born_step_radix_orders_match_oracle(shape, C, kind, dtype, order, device):
    """One Born step on the LDS-staged kernels for every radix order, against the oracle."""
    import torch
    print(f'born radix order {order} {shape} C={C} chi_kind={kind} {dtype}', flush=True)
    E, s, chi, alphas, a, c = _born_case(shape, C, kind, seed=order + 5)
    step = [0.41 + 0.03 * i for i in range(len(shape))]
    k = ob.k_vectors(shape, step)
    tdt = torch.complex64 if dtype == 'c64' else torch.complex128
    rdt = torch.float32 if dtype == 'c64' else torch.float64
    tol = 1e-5 if dtype == 'c64' else 1e-11
    plan = _plan(shape, step, tdt, device)
    plan.set_param('reg_fft', 0)
    plan.set_param('radix_order', order)
    Et = torch.as_tensor(E[0], dtype=tdt, device=device).contiguous()
    Eo = torch.zeros_like(Et)
    st = torch.as_tensor(s[0], dtype=tdt, device=device).contiguous()
    ct = torch.as_tensor(_pack(chi[0], kind, np.complex64 if tdt == torch.complex64 else np.complex128),
                         device=device).contiguous()
    coef = torch.tensor([alphas[0].real, alphas[0].imag, a[0], c[0].real, c[0].imag, 0.0], dtype=rdt, device=device)
    norms = torch.zeros(2, dtype=torch.float64, device=device)
    plan.born_step(Et, Eo, st, ct, C, kind, coef, 1, 0, 0, norms)
    torch.cuda.synchronize()
    d_ref = _oracle_update(E[0], s[0], chi[0], alphas[0], a[0], c[0], k, C == 3)
    assert rel_l2(Eo.cpu().numpy() - E[0], d_ref) < tol


@pytest.mark.parametrize('shape,C,kind,shared', [((64, 48), 1, 1, True), ((12, 16, 20), 3, 2, True),
                                                 ((12, 16, 20), 3, 1, False), ((512,), 3, 2, False),
                                                 ((12, 16, 20), 3, 3, False), ((12, 16, 20), 3, 4, False),
                                                 ((64, 32, 64), 3, 4, False), ((13, 17, 11), 3, 2, False),
                                                ((101, 24), 1, 1, True)],
                         ids=lambda v: str(v))
def test_born_step_batched(shape, C, kind, shared, device):
    import torch
    print(f'born batched {shape} C={C} kind={kind} shared={shared}', flush=True)
    nb = 3
    E, s, chi, alphas, a, c = _born_case(shape, C, kind, nbatch=nb, seed=11, chi_shared=shared)
    step = [0.5] * len(shape)
    k = ob.k_vectors(shape, step)
    plan = _plan(shape, step, torch.complex64, device)
    Et = torch.as_tensor(E, dtype=torch.complex64, device=device).contiguous()
    Eo = torch.zeros_like(Et)
    st = torch.as_tensor(s, dtype=torch.complex64, device=device).contiguous()
    ct = torch.as_tensor(np.stack([_pack(_, kind, np.complex64) for _ in chi]), device=device).contiguous()
    coef = torch.tensor([[al.real, al.imag, aa, cc.real, cc.imag, 0.0] for al, aa, cc in zip(alphas, a, c)],
                        dtype=torch.float32, device=device)
    norms = torch.zeros(2 * nb, dtype=torch.float64, device=device)
    V = int(np.prod(shape))
    chi_size = {1: 1, 2: 9, 3: 6, 4: 4}[kind] * V
    plan.born_step(Et, Eo, st, ct, C, kind, coef, nb, C * V, 0 if shared else chi_size, norms)
    torch.cuda.synchronize()
    out = Eo.cpu().numpy()
    nrm = norms.cpu().numpy()
    for b in range(nb):
        d_ref = _oracle_update(E[b], s[b], chi[0 if shared else b], alphas[b], a[b], c[b], k, C == 3)
        assert rel_l2(out[b] - E[b], d_ref) < 1e-5
        assert abs(nrm[2 * b] - np.sum(np.abs(d_ref) ** 2)) / np.sum(np.abs(d_ref) ** 2) < 2e-5


@pytest.mark.parametrize('shape', [(5, 10, 20), (32, 24, 16), (40, 48), (64,)], ids=lambda s: 'x'.join(map(str, s)))
def test_longitudinal_projection_and_chi_mul(shape, device):
    import torch
    print(f'lproj/chi_mul {shape}', flush=True)
    rng = np.random.default_rng(3)
    F = rng.standard_normal((3, 1, *shape)) + 1j * rng.standard_normal((3, 1, *shape))
    k = ob.k_vectors(shape, [1.0] * len(shape))
    plan = _plan(shape, [1.0] * len(shape), torch.complex64, device)
    Ft = torch.as_tensor(F, dtype=torch.complex64, device=device).contiguous()
    out = torch.empty_like(Ft)
    plan.longitudinal_projection(Ft, out)
    assert rel_l2(out.cpu().numpy(), ob.longitudinal_projection_ft(F, k)) < 1e-6
    chi = rng.standard_normal((3, 3, *shape)) + 1j * rng.standard_normal((3, 3, *shape))
    ct = torch.as_tensor(chi, dtype=torch.complex64, device=device).contiguous()
    y = torch.empty_like(Ft)
    plan.chi_mul(ct, 2, Ft, y, 3, 1)
    assert rel_l2(y.cpu().numpy(), ob.chi_mul(chi, F)) < 1e-6
    x = torch.empty(1, dtype=torch.float64, device=device)
    plan.norm2(Ft, x)
    F32 = F.astype(np.complex64).astype(np.complex128)  # the values the device holds, summed in float64
    assert abs(float(x.item()) - np.sum(np.abs(F32) ** 2)) / np.sum(np.abs(F32) ** 2) < 1e-12


@pytest.mark.parametrize('shape,C,kind,dtype', [((256, 256, 256), 3, 3, 'c64'), ((64, 64, 64), 3, 2, 'c64'),
                                                ((96, 40, 24), 3, 1, 'c64'), ((128, 96), 1, 1, 'c64'),
                                                ((1024,), 3, 2, 'c64'), ((512,), 3, 3, 'c64'),
                                                ((64, 128, 64), 3, 1, 'c64'), ((128, 64, 256), 3, 3, 'c64'),
                                                ((64, 64, 128), 1, 1, 'c64'), ((256, 64), 3, 2, 'c64'),
                                                ((16, 20, 24), 3, 3, 'c128'), ((48, 64), 1, 1, 'c128'),
                                                ((256, 256, 256), 3, 4, 'c64'), ((64, 128, 64), 3, 4, 'c64'),
                                                ((16, 20, 24), 3, 4, 'c128'), ((96, 96, 96), 3, 4, 'c64'),
                                                ((192, 384), 1, 1, 'c64'), ((64, 96, 384), 3, 3, 'c64'),
                                                ((80, 80, 80), 3, 4, 'c64'), ((160, 320), 1, 1, 'c64'),
                                                ((1024, 256), 3, 3, 'c64'), ((768, 640), 1, 1, 'c64'),
                                                ((64, 2048), 3, 4, 'c64'), ((32, 2048), 3, 2, 'c64'),
                                                ((2048,), 3, 1, 'c64'), ((2048,), 1, 1, 'c64'),
                                                ((13, 17, 11), 3, 3, 'c64'), ((13, 17, 11), 3, 2, 'c128'),
                                                ((101, 64), 1, 1, 'c64'), ((8192,), 3, 1, 'c128'),
                                                ((4096,), 3, 4, 'c128')],
                         ids=lambda v: str(v))
def test_prologue_iterate_matches_step(shape, C, kind, dtype, device):
    """mmx_born_prologue + mmx_born_iterate == mmx_born_step, and W afterwards == prologue(Eout)."""
    import torch
    print(f'prologue/iterate {shape} C={C} kind={kind} {dtype}', flush=True)
    E, s, chi, alphas, a, c = _born_case(shape, C, kind, seed=5)
    step = [0.41] * len(shape)
    tdt = torch.complex64 if dtype == 'c64' else torch.complex128
    rdt = torch.float32 if dtype == 'c64' else torch.float64
    tol = 2e-6 if dtype == 'c64' else 1e-12
    plan = _plan(shape, step, tdt, device)
    Et = torch.as_tensor(E[0], dtype=tdt, device=device).contiguous()
    st = torch.as_tensor(s[0], dtype=tdt, device=device).contiguous()
    ct = torch.as_tensor(_pack(chi[0], kind, np.complex64 if tdt == torch.complex64 else np.complex128),
                         device=device).contiguous()
    coef = torch.tensor([alphas[0].real, alphas[0].imag, a[0], c[0].real, c[0].imag, 0.0], dtype=rdt, device=device)
    n1 = torch.zeros(2, dtype=torch.float64, device=device)
    n2 = torch.zeros(2, dtype=torch.float64, device=device)
    E1 = torch.empty_like(Et)
    plan.born_step(Et, E1, st, ct, C, kind, coef, 1, 0, 0, n1)
    W = torch.empty_like(Et)
    E2 = torch.empty_like(Et)
    plan.born_prologue(Et, W, st, ct, C, kind, coef)
    plan.born_iterate(Et, E2, W, st, ct, C, kind, coef, n2)
    W_next = torch.empty_like(Et)
    plan.born_prologue(E2, W_next, st, ct, C, kind, coef)
    torch.cuda.synchronize()
    d1 = (E1 - Et)
    assert float(torch.linalg.norm(E2 - E1) / torch.linalg.norm(d1)) < tol
    assert float(((n2 - n1).abs() / n1.abs()).max()) < 10 * tol
    assert float(torch.linalg.norm(W - W_next) / torch.linalg.norm(W_next)) < tol


@pytest.mark.parametrize('shape,C,kind,dtype', [((128, 128, 128), 3, 4, 'c64'), ((64, 96), 1, 1, 'c64'),
                                                ((512,), 3, 3, 'c64'), ((13, 17, 11), 3, 3, 'c64'),
                                                ((16, 20, 24), 3, 2, 'c128')],
                         ids=lambda v: str(v))
def test_iterate_split_matches_iterate(shape, C, kind, dtype, device):
    """mmx_born_iterate_split (norm reduction on a second stream, partial slots 0 and 1 alternating) gives the
    fields, work buffer and norms of mmx_born_iterate, bit for bit, over four iterations. As its contract asks,
    call n + 2 is issued only after the reduction of call n has run (an event on the norm stream), as
    Solution.step does."""
    import torch
    print(f'iterate split {shape} C={C} kind={kind} {dtype}', flush=True)
    E, s, chi, alphas, a, c = _born_case(shape, C, kind, seed=7)
    tdt = torch.complex64 if dtype == 'c64' else torch.complex128
    rdt = torch.float32 if dtype == 'c64' else torch.float64
    plan = _plan(shape, [0.37] * len(shape), tdt, device)
    Et = torch.as_tensor(E[0], dtype=tdt, device=device).contiguous()
    st = torch.as_tensor(s[0], dtype=tdt, device=device).contiguous()
    ct = torch.as_tensor(_pack(chi[0], kind, np.complex64 if tdt == torch.complex64 else np.complex128),
                         device=device).contiguous()
    coef = torch.tensor([alphas[0].real, alphas[0].imag, a[0], c[0].real, c[0].imag, 0.0], dtype=rdt, device=device)
    norm_stream = torch.cuda.Stream(device=device)
    runs = {}
    for split in (False, True):
        bufs = [Et.clone(), torch.empty_like(Et), torch.empty_like(Et)]
        W = torch.empty_like(Et)
        plan.born_prologue(bufs[0], W, st, ct, C, kind, coef)
        norms = [torch.full((2,), -1.0, dtype=torch.float64, device=device) for _ in range(4)]
        hosts = [torch.full((2,), -2.0, dtype=torch.float64).pin_memory() for _ in range(4)]
        reduced = []
        for i in range(4):
            src, dst = bufs[i % 3], bufs[(i + 1) % 3]
            if split:
                if i >= 2:
                    reduced[i - 2].synchronize()
                plan.born_iterate(src, dst, W, st, ct, C, kind, coef, norms[i], norm_stream=norm_stream, slot=i % 2,
                                  norms_host=hosts[i] if i % 2 else None)
                ev = torch.cuda.Event()
                ev.record(norm_stream)
                reduced.append(ev)
            else:
                plan.born_iterate(src, dst, W, st, ct, C, kind, coef, norms[i])
        torch.cuda.synchronize()
        runs[split] = ([b.cpu().numpy() for b in bufs], W.cpu().numpy(), [n.cpu().numpy() for n in norms])
        if split:  # the host copies (odd calls) equal the device norms; no copy was asked for the even ones
            for i in range(4):
                assert np.array_equal(hosts[i].numpy(), runs[True][2][i] if i % 2 else np.full(2, -2.0))
    for b0, b1 in zip(runs[False][0], runs[True][0]):
        assert np.array_equal(b0, b1)
    assert np.array_equal(runs[False][1], runs[True][1])
    for n0, n1 in zip(runs[False][2], runs[True][2]):
        assert np.array_equal(n0, n1) and np.all(n0 > 0)


def test_iterate_split_batch_matches_iterate(device):
    """mmx_born_iterate_split on a batch of three problems (own alpha and coefficients each, shared chi) gives the
    fields and per-problem norms of mmx_born_iterate bit for bit; the host copy holds all 2 x 3 norms."""
    import torch
    shape, C, kind, nb = (64, 40, 32), 3, 4, 3
    print(f'iterate split batch {shape} nbatch={nb}', flush=True)
    E, s, chi, alphas, a, c = _born_case(shape, C, kind, nbatch=nb, seed=9)
    plan = _plan(shape, [0.37] * 3, torch.complex64, device)
    Et = torch.as_tensor(E, dtype=torch.complex64, device=device).contiguous()
    st = torch.as_tensor(s, dtype=torch.complex64, device=device).contiguous()
    ct = torch.as_tensor(_pack(chi[0], kind, np.complex64), device=device).contiguous()
    coef = torch.tensor([[al.real, al.imag, a[b], c[b].real, c[b].imag, 0.0] for b, al in enumerate(alphas)],
                        dtype=torch.float32, device=device)
    norm_stream = torch.cuda.Stream(device=device)
    out = {}
    for split in (False, True):
        W = torch.empty_like(Et)
        Eo = torch.empty_like(Et)
        norms = torch.full((2 * nb,), -1.0, dtype=torch.float64, device=device)
        host = torch.full((2 * nb,), -2.0, dtype=torch.float64).pin_memory()
        plan.born_prologue(Et, W, st, ct, C, kind, coef, nb, C * int(np.prod(shape)), 0)
        if split:
            plan.born_iterate(Et, Eo, W, st, ct, C, kind, coef, norms, nb, C * int(np.prod(shape)), 0,
                              norm_stream=norm_stream, slot=1, norms_host=host)
        else:
            plan.born_iterate(Et, Eo, W, st, ct, C, kind, coef, norms, nb, C * int(np.prod(shape)), 0)
        torch.cuda.synchronize()
        out[split] = (Eo.cpu().numpy(), W.cpu().numpy(), norms.cpu().numpy(), host.numpy().copy())
    assert np.array_equal(out[True][0], out[False][0]) and np.array_equal(out[True][1], out[False][1])
    assert np.array_equal(out[True][2], out[False][2]) and np.all(out[False][2] > 0)
    assert np.array_equal(out[True][3], out[False][2])


def test_iterate_split_rejects_bad_arguments(device):
    """mmx_born_iterate_split refuses a slot other than 0 and 1 and a norm stream equal to the launch stream."""
    import torch
    from macromax_amd import _native
    shape = (32, 32)
    plan = _plan(shape, [0.5, 0.5], torch.complex64, device)
    z = [torch.zeros(shape, dtype=torch.complex64, device=device) for _ in range(5)]
    coef = torch.tensor([1.0, 1.0, 1.0, 0.0, 0.0, 0.0], dtype=torch.float32, device=device)
    norms = torch.zeros(2, dtype=torch.float64, device=device)
    with pytest.raises(_native.NativeError, match='slot'):
        plan.born_iterate(z[0], z[1], z[2], z[3], z[4], 1, 1, coef, norms, norm_stream=torch.cuda.Stream(device=device),
                          slot=2)
    launch = torch.cuda.Stream(device=device)
    with torch.cuda.stream(launch), pytest.raises(_native.NativeError, match='norm_stream'):
        plan.born_iterate(z[0], z[1], z[2], z[3], z[4], 1, 1, coef, norms, norm_stream=launch, slot=0)
    torch.cuda.synchronize()


@pytest.mark.parametrize('shape,C,kind', [((64, 128, 64), 3, 3), ((20, 24, 40), 3, 2), ((48, 40), 1, 1),
                                          ((256,), 3, 1), ((256, 64), 3, 1), ((64, 128, 64), 3, 4)],
                         ids=lambda v: str(v))
def test_source_box(shape, C, kind, device):
    """A source zero outside a box: the box-restricted passes (born_step, prologue/iterate) give the same update as
    the dense ones; a box that cuts the source off changes it (the box is honoured)."""
    import torch
    print(f'source box {shape} C={C} kind={kind}', flush=True)
    E, s, chi, alphas, a, c = _born_case(shape, C, kind, seed=11)
    nd = len(shape)
    lo = [max(0, n // 4) for n in shape]
    hi = [min(n, n // 4 + max(1, n // 3)) for n in shape]
    box = tuple(slice(l, h) for l, h in zip(lo, hi))
    s0 = np.zeros_like(s[0])
    s0[(slice(None), slice(None)) + box] = s[0][(slice(None), slice(None)) + box]
    step = [0.41] * nd
    plan = _plan(shape, step, torch.complex64, device)
    Et = torch.as_tensor(E[0], dtype=torch.complex64, device=device).contiguous()
    st = torch.as_tensor(s0, dtype=torch.complex64, device=device).contiguous()
    ct = torch.as_tensor(_pack(chi[0], kind, np.complex64), device=device).contiguous()
    coef = torch.tensor([alphas[0].real, alphas[0].imag, a[0], c[0].real, c[0].imag, 0.0], dtype=torch.float32,
                        device=device)
    out = {}
    for name, sb in (('dense', None), ('box', (lo, hi)), ('cut', (lo, [l + 1 for l in lo]))):
        n1 = torch.zeros(2, dtype=torch.float64, device=device)
        E1 = torch.empty_like(Et)
        plan.born_step(Et, E1, st, ct, C, kind, coef, 1, 0, 0, n1, s_box=sb)
        W = torch.empty_like(Et)
        E2 = torch.empty_like(Et)
        n2 = torch.zeros(2, dtype=torch.float64, device=device)
        plan.born_prologue(Et, W, st, ct, C, kind, coef, s_box=sb)
        plan.born_iterate(Et, E2, W, st, ct, C, kind, coef, n2, s_box=sb)
        torch.cuda.synchronize()
        out[name] = (E1.cpu().numpy(), E2.cpu().numpy(), W.cpu().numpy())
    for k in range(3):
        assert np.array_equal(out['box'][k], out['dense'][k]), k
    assert not np.array_equal(out['cut'][0], out['dense'][0])
