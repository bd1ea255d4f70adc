"""ScatteringMatrix and DepositionMatrix on the HIP backend: the reference's known answers
(python/tests/test_matrix.py:16-381, 443-485) on the same grids, bounds and media, plus the batched column solve
against the reference's column-by-column loop.

Known answers: free-space plane waves of unit-norm vectors (analytic fields; decimal 7), the plane source producing
the free-space field (decimal 6), solved plane waves matching free space away from the boundaries (decimal 2),
detection of an outward free-space field returning its input vector (decimal 8), time reversal of a detection
vector, the empty-space matrix being the identity (decimal 2), lossless thin films and random layers giving unitary
matrices (decimal 2), quarters assembling the matrix exactly and small inverse problems on the forward transmission
(decimal 6 and 3). The batched solve of all columns equals the per-column Solution loop to 1e-9 relative (complex128).
"""
import numpy as np
import numpy.testing as npt
import pytest

pytestmark = pytest.mark.gpu

WAVELENGTH = 500e-9
K0 = 2 * np.pi / WAVELENGTH


def _cases():
    from macromax_amd import Grid
    from macromax_amd.bound import LinearBound
    grids = (Grid([256, 7], WAVELENGTH / 4), Grid([256, 5, 3], WAVELENGTH / 4),
             Grid([256, 8], [WAVELENGTH / 4, WAVELENGTH / 5]), Grid([256], WAVELENGTH / 4))
    out = []
    for g in grids:
        thickness = np.zeros([g.ndim, 1])
        thickness[0, 0] = 20 * WAVELENGTH  # periodic in the transverse directions
        out.append((g, LinearBound(g, thickness=thickness, max_extinction_coefficient=0.25)))
    return out


CASE_IDS = ['2d7', '3d5x3', '2d8aniso', '1d']
PARAMS = [(v, i) for v in (False, True) for i in range(4)]
PARAM_IDS = [f'{"vec" if v else "scal"}-{CASE_IDS[i]}' for v, i in PARAMS]


def _matrix(vectorial, idx, n=1, **kwargs):
    from macromax_amd import matrix
    grid, bound = _cases()[idx]
    return matrix.ScatteringMatrix(grid, vectorial=vectorial, vacuum_wavelength=WAVELENGTH, refractive_index=n,
                                   bound=bound, **kwargs), grid, bound


def _mode_k(grid, m, angle_idx):
    """Unit k-vector of propagating mode angle_idx (raster order over the centred transverse k grid)."""
    kt = np.zeros([2, *grid.shape[1:]])
    for a in range(1, grid.ndim):
        kt[a - 1] = np.fft.fftshift(np.asarray(grid.k[a]).ravel()).reshape(
            [-1 if b == a - 1 else 1 for b in range(grid.ndim - 1)]) / K0
    prop = np.sqrt(np.sum(kt ** 2, axis=0)) + np.sqrt(np.finfo(m.dtype).eps) < 1
    kt = kt[:, prop][:, angle_idx]
    return np.array([np.sqrt(1 - np.sum(kt ** 2)), *kt]), kt


@pytest.mark.parametrize('vectorial,idx', PARAMS, ids=PARAM_IDS)
def test_srcvec2freespace(vectorial, idx):
    print(f'srcvec2freespace {PARAM_IDS[PARAMS.index((vectorial, idx))]}', flush=True)
    m, grid, _ = _matrix(vectorial, idx)
    for j in range(m.shape[1]):
        v = np.zeros(m.shape[1], dtype=np.complex128)
        v[j] = 1 + 1j
        side = j // (v.size // 2)
        angle = j % (v.size // 2)
        if vectorial:
            pol = np.zeros(2)
            pol[angle % 2] = 1
            angle //= 2
        kdir, kt = _mode_k(grid, m, angle)
        fld = m.srcvec2freespace(v)
        assert fld.shape == (1 + 2 * vectorial, *grid.shape)
        centre = fld.reshape(fld.shape[0], -1)[:, np.ravel_multi_index(np.array(fld.shape[1:]) // 2, fld.shape[1:])]
        npt.assert_almost_equal(np.linalg.norm(centre), np.linalg.norm(v) / np.sqrt(kdir[0]))
        if vectorial:
            rn = np.linalg.norm(kt)
            rdir = kt / (rn + (rn == 0))
            rpol = rdir @ pol
            azim = np.array([0, *(pol - rdir * rpol)])
            radial = np.array([-np.sqrt(1 - kdir[0] ** 2) * rpol, *(rdir * rpol * kdir[0])])
            expected = v[j] * (radial + azim)
        else:
            expected = np.atleast_1d(v[j])
        expected = expected / np.sqrt(kdir[0])
        npt.assert_almost_equal(centre, expected)
        expected = expected.reshape(-1, *([1] * grid.ndim)) * np.exp(
            (1 - 2 * side) * 1j * K0 * sum(grid[a] * kdir[a] for a in range(grid.ndim)))
        npt.assert_array_almost_equal(fld, expected)


@pytest.mark.parametrize('vectorial,idx', PARAMS, ids=PARAM_IDS)
def test_srcvec2source(vectorial, idx):
    """The plane source convolved with the 1D Green's function of each transverse mode gives the free-space field
    on its plane (test_matrix.py:106-127)."""
    print(f'srcvec2source {PARAM_IDS[PARAMS.index((vectorial, idx))]}', flush=True)
    m, grid, _ = _matrix(vectorial, idx)
    axes = tuple(1 + np.arange(grid.ndim))
    for j in range(m.shape[1]):
        v = np.zeros(m.shape[1], dtype=np.complex128)
        v[j] = 1 + 1j
        free = m.srcvec2freespace(v)
        src = m.srcvec2source(v)
        intensity = np.sum(np.sum(np.abs(src) ** 2, axis=0).reshape(src.shape[1], -1), -1)
        planes = np.flatnonzero(~np.isclose(intensity, 0.0))
        assert planes.size == 1
        fwd = np.sqrt(np.maximum(1.0 - sum((np.asarray(grid.k[a]) / K0) ** 2 for a in range(1, grid.ndim)), 0.0))
        s2f = 1j * grid.step[0] / (2 * K0 * fwd + np.isclose(fwd, 0.0))
        from_src = np.fft.ifftn(np.fft.fftn(src, axes=axes) * s2f, axes=axes)
        npt.assert_array_almost_equal(from_src[:, planes], free[:, planes])


@pytest.mark.parametrize('vectorial,idx', PARAMS, ids=PARAM_IDS)
def test_srcvec2detfield(vectorial, idx):
    print(f'srcvec2detfield {PARAM_IDS[PARAMS.index((vectorial, idx))]}', flush=True)
    m, grid, bound = _matrix(vectorial, idx)
    x = np.asarray(grid[0]).ravel()
    free_idx = np.flatnonzero((grid.first[0] + bound.thickness[0, 0] + 4 * WAVELENGTH < x)
                              & (x < x[-1] - bound.thickness[0, -1] - 4 * WAVELENGTH))
    assert free_idx.size > 2
    for j in range(m.shape[1]):
        v = np.zeros(m.shape[1], dtype=np.complex128)
        v[j] = 1 + 1j
        free = m.srcvec2freespace(v)
        fld = m.srcvec2detfield(v)
        npt.assert_array_almost_equal(np.abs(fld[:, free_idx]), np.abs(free[:, free_idx]), decimal=2)
        npt.assert_array_almost_equal(fld[:, free_idx], free[:, free_idx], decimal=2)


@pytest.mark.parametrize('vectorial,idx', PARAMS, ids=PARAM_IDS)
def test_detfield2detvec(vectorial, idx):
    print(f'detfield2detvec {PARAM_IDS[PARAMS.index((vectorial, idx))]}', flush=True)
    m, grid, _ = _matrix(vectorial, idx)
    for j in range(m.shape[1]):
        v = np.zeros(m.shape[1], dtype=np.complex128)
        v[j] = 1 + 1j
        outward = 0.0
        for side in range(2):
            vmask = np.zeros_like(v)
            if side == 0:
                vmask[:v.size // 2] = 1
            else:
                vmask[v.size // 2:] = 1
            free = m.srcvec2freespace(v * vmask)
            omask = np.zeros(free.shape[1:])
            if side == 0:
                omask[omask.shape[0] // 2:] = 1
            else:
                omask[:omask.shape[0] // 2] = 1
            outward = outward + free * omask
        out = m.detfield2detvec(outward)
        npt.assert_array_almost_equal(np.linalg.norm(out), np.linalg.norm(v), decimal=8)
        npt.assert_array_almost_equal(out, v, decimal=8)


@pytest.mark.parametrize('vectorial,idx', PARAMS, ids=PARAM_IDS)
def test_detvec2srcvec(vectorial, idx):
    print(f'detvec2srcvec {PARAM_IDS[PARAMS.index((vectorial, idx))]}', flush=True)
    m, grid, _ = _matrix(vectorial, idx)
    k = np.array([1, 0, 0]) * K0
    incident = np.exp(1j * sum(-np.abs(r * kk) for r, kk in zip(grid, k)))[np.newaxis]
    emanating = incident.conj()
    for side_mask in (np.asarray(grid[0]) < 0, np.asarray(grid[0]) > 0):
        pol = np.array([0, 1, 0]) if vectorial else np.array([1])
        pol = pol.reshape(-1, *([1] * grid.ndim))
        emanating = pol * emanating
        detvec = m.detfield2detvec(emanating * side_mask)
        reversed_fld = m.srcvec2freespace(m.detvec2srcvec(detvec))
        sel = np.broadcast_to(side_mask, grid.shape).ravel()
        npt.assert_array_almost_equal(reversed_fld.reshape(reversed_fld.shape[0], -1)[:, sel].conj(),
                                      np.broadcast_to(emanating, reversed_fld.shape).reshape(reversed_fld.shape[0], -1)[:, sel])


@pytest.mark.parametrize('vectorial,idx', PARAMS, ids=PARAM_IDS)
def test_matmul_and_identity(vectorial, idx):
    """Products with zero and basis vectors in free space (test_matrix.py:212-226) and the full empty-space matrix
    (test_matrix.py:228-241)."""
    print(f'matmul {PARAM_IDS[PARAMS.index((vectorial, idx))]}', flush=True)
    m, grid, _ = _matrix(vectorial, idx)
    z = np.zeros(m.shape[1], dtype=m.dtype)
    npt.assert_array_equal(m @ z, z)
    for j in range(m.shape[1]):
        v = np.zeros(m.shape[1], dtype=m.dtype)
        v[j] = 1 + 1j
        p = m @ v
        npt.assert_almost_equal(np.linalg.norm(p), np.linalg.norm(v), decimal=2)
        npt.assert_array_almost_equal(p, v, decimal=2)
    m2, _, _ = _matrix(vectorial, idx, n=None)
    npt.assert_array_almost_equal(np.abs(m2), np.eye(*m2.shape), decimal=2)
    npt.assert_array_almost_equal(m2, np.eye(*m2.shape), decimal=2)


def _thin_film(grid):
    n = np.ones(grid.shape)
    c = grid.shape[0] // 2
    for off in (-11, -10, -9, -8, -1, 0, 2):
        n[c + off] = 1.5
    return n


def _random_layer(grid):
    rng = np.random.Generator(np.random.PCG64(seed=1))
    layer = np.array(grid.shape)
    layer[0] = 8 * WAVELENGTH // grid.step[0]
    mat = np.zeros(grid.shape, dtype=bool)
    while np.mean(mat) < 0.75:
        centre = rng.random(grid.ndim) * grid.extent + grid.first
        radius = rng.uniform(0.75, 1.5) * WAVELENGTH / 2
        mat = np.logical_or(mat, sum((x - c) ** 2 for x, c in zip(grid, centre)) <= radius ** 2)
    mat = 1 + 0.5 * mat
    n = np.ones(grid.shape)
    rows = grid.shape[0] // 2 + np.arange(layer[0]) - layer[0] // 2
    n[rows] = mat[rows]
    return n


@pytest.mark.parametrize('vectorial,idx', PARAMS, ids=PARAM_IDS)
def test_unitarity_and_quadrants(vectorial, idx):
    """Lossless media scatter unitarily (test_matrix.py:243-293); the quarters assemble the matrix and support
    the inverse problem on the forward transmission (test_matrix.py:295-381)."""
    from macromax_amd import matrix
    print(f'unitarity {PARAM_IDS[PARAMS.index((vectorial, idx))]}', flush=True)
    grid, _ = _cases()[idx]
    for kind, n in (('film', _thin_film(grid)), ('random', _random_layer(grid))):
        m, _, _ = _matrix(vectorial, idx, n=n)
        arr = np.asarray(m)
        npt.assert_almost_equal(np.sqrt(np.sum(np.abs(arr) ** 2, axis=0)), 1.0, decimal=2, err_msg=kind)
        npt.assert_almost_equal(np.linalg.svd(arr, compute_uv=False), 1.0, decimal=2, err_msg=kind)
        m_ft, m_fr, m_br, m_bt = m.forward_transmission, m.front_reflection, m.back_reflection, m.backward_transmission
        npt.assert_array_equal(m, np.vstack([np.hstack([m_ft, m_br]), np.hstack([m_fr, m_bt])]))
        target = np.zeros(m_ft.shape[0], dtype=np.complex128)
        target[0] = 0.5 + 0.5j if kind == 'film' else 1.0
        source = matrix.inv(m_ft, 1e-6) @ target
        npt.assert_array_almost_equal(m_ft @ source, target, decimal=6 if kind == 'film' else 3)
        if kind == 'random':
            w, vecs = np.linalg.eig(m_ft)
            top = int(np.argmax(np.abs(w)))
            npt.assert_array_almost_equal(m_ft @ vecs[:, top], w[top] * vecs[:, top], decimal=3)


@pytest.mark.parametrize('vectorial', [False, True], ids=['scal', 'vec'])
def test_batched_columns_match_sequential_loop(vectorial):
    """All columns from one batched solve equal the reference's column-by-column Solution loop (same stop rule)."""
    print(f'batched vs sequential columns {"vec" if vectorial else "scal"}', flush=True)
    grid, _ = _cases()[1 if vectorial else 0]
    n = _random_layer(grid)
    m_b, _, _ = _matrix(vectorial, 1 if vectorial else 0, n=n)
    m_s, _, _ = _matrix(vectorial, 1 if vectorial else 0, n=n,
                        callback=lambda s: s.iteration < 1e4 and s.residue > 1e-6)
    a_b = np.asarray(m_b)
    a_s = np.asarray(m_s)
    assert np.linalg.norm(a_b - a_s) <= 1e-9 * np.linalg.norm(a_s)
    # cached columns scale with the input; non-basis products solve their own column
    v = np.linspace(0.5, 1.5, m_b.shape[1]) * (1 + 0.5j)
    p = m_b @ v  # a new solve of the superposition, to the same residue 1e-6 as the columns
    assert np.linalg.norm(p - a_s @ v) <= 1e-4 * np.linalg.norm(a_s @ v)
    npt.assert_array_almost_equal(m_b @ (3j * np.eye(m_b.shape[1])[:, 2]), 3j * a_s[:, 2], decimal=10)


@pytest.mark.parametrize('vectorial,idx', [(False, 0), (True, 1)], ids=['scal-2d7', 'vec-3d5x3'])
def test_deposition_matrix(vectorial, idx):
    """Shapes and dtypes of deposition matrices (test_matrix.py:443-485); the default operators reproduce the
    scattering matrix without the incident-field subtraction, checked on the empty-space matrix."""
    from macromax_amd import matrix
    print(f'deposition {"vec" if vectorial else "scal"}', flush=True)
    s, grid, _ = _matrix(vectorial, idx)
    dep = matrix.DepositionMatrix(s)
    assert dep.dtype == s.dtype and dep.shape == s.shape
    assert (dep @ np.ones(dep.shape[1])).shape == (s.shape[0],)
    assert (dep @ np.ones([dep.shape[1], 1])).shape == (s.shape[0], 1)
    assert (dep @ np.ones([dep.shape[1], 2])).shape == (s.shape[0], 2)
    full = dep @ np.eye(dep.shape[0])
    assert full.shape == s.shape

    def centre(fld):
        sel = np.asarray(fld).reshape([1 + 2 * vectorial, *s.grid.shape])
        while sel.ndim > 0:
            sel = sel[sel.shape[0] // 2]
        return sel
    d1 = matrix.DepositionMatrix(s, output_operator=centre)
    assert d1.dtype == s.dtype and d1.shape == (1, s.shape[1])
    d2 = matrix.DepositionMatrix(s, output_operator=lambda f: np.asarray([f[0], f[-1]]))
    assert d2.shape == (2, s.shape[1])
    d_all = matrix.DepositionMatrix(s, output_operator=lambda f: f)
    assert d_all.shape == ((1 + 2 * vectorial) * s.grid.size, s.shape[1])
    # the field at the centre for basis input j is the free-space plane wave of mode j there
    row = np.asarray(d1 @ np.eye(s.shape[1])).ravel()
    expected = np.array([centre(s.srcvec2freespace(np.eye(s.shape[1])[j])) for j in range(s.shape[1])])
    npt.assert_array_almost_equal(row, expected, decimal=2)


def test_matrix_api_edges():
    """Caching on/off, a precomputed array, element access, out= arguments and mismatched shapes."""
    from macromax_amd import matrix
    print('matrix api edges', flush=True)
    m, grid, _ = _matrix(False, 0, n=_thin_film(_cases()[0][0]))
    arr = np.asarray(m)
    # element access computes only the selected columns; rows come from the full matrix
    npt.assert_array_almost_equal(m[1, 2], arr[1, 2], decimal=12)
    npt.assert_array_almost_equal(m[:, 3], arr[:, 3], decimal=12)
    npt.assert_array_almost_equal(m[2], arr[2], decimal=12)
    # no caching: every product solves (same numbers to the solver tolerance)
    m2, _, _ = _matrix(False, 0, n=_thin_film(grid), caching=False)
    assert not m2.caching
    v = np.zeros(m2.shape[1], dtype=np.complex128)
    v[1] = 2.0
    npt.assert_array_almost_equal(m2 @ v, 2.0 * arr[:, 1], decimal=8)
    # a precomputed array pre-fills the cache: products need no solve
    m3, _, _ = _matrix(False, 0, n=_thin_film(grid), array=arr)
    npt.assert_array_equal(m3 @ v, 2.0 * arr[:, 1])
    with pytest.raises(ValueError):
        _matrix(False, 0, array=np.zeros((3, 3)))
    with pytest.raises(ValueError):  # scipy's LinearOperator checks the shape first
        m @ np.zeros(m.shape[1] + 1)
    with pytest.raises(KeyError):
        m[0, 0] = 1.0
    # out= arguments
    src = np.empty((1, *grid.shape), dtype=m.dtype)
    res = m.srcvec2source(v, out=src)
    assert res is src and np.any(src != 0)
    fld = np.empty((1, *grid.shape), dtype=m.dtype)
    res = m.srcvec2detfield(v, out=fld)
    assert res is fld
    # v enters from the front: its incident wave is removed at the front detection plane only
    back_det, front_det = m.detector_planes
    out = fld.copy()
    out[:, front_det] -= m.srcvec2freespace(v)[:, front_det]
    npt.assert_array_almost_equal(m.detfield2detvec(out), m @ v, decimal=6)
    # the transfer matrix converts back
    npt.assert_array_almost_equal(matrix.convert(m.transfer()), arr, decimal=8)
