"""macromax_amd.grid (Grid, MutableGrid) against the known answers of the reference's Grid tests
(python/tests/test_grid.py:10-363), as tables of constructor arguments and expected coordinates (CPU)."""
import numpy as np
import numpy.testing as npt
import pytest

from macromax_amd.grid import Grid, MutableGrid

# (args, kwargs, expected coordinates of a one-dimensional grid)
COORDINATES = [
    # shape from extent, first/last/center and step (test_grid.py:10-36)
    ((), dict(extent=10, step=1), np.arange(-5, 5)),
    ((), dict(extent=11, step=1), np.arange(-5, 6)),
    ((), dict(extent=10, step=2), np.arange(-4, 6, 2)),
    ((), dict(extent=11, step=2), np.arange(-6, 5, 2)),
    ((), dict(extent=10, step=3), np.arange(-6, 4, 3)),
    ((), dict(extent=11, step=3), np.arange(-6, 6, 3)),
    ((), dict(extent=11, step=-3), np.arange(6, -6, -3)),
    ((), dict(extent=12, step=3), np.arange(-6, 6, 3)),
    ((), dict(first=-5, last=5, step=1), np.arange(-5, 5)),
    ((), dict(first=-5, center=0, step=1), np.arange(-5, 5)),
    ((), dict(center=0, last=5, step=1), np.arange(-5, 5)),
    ((), dict(first=-2, last=3, step=1), np.arange(-2, 3)),
    ((), dict(first=-1.5, center=0.5, step=1), np.arange(-1.5, 2.5)),
    # unit step (:38-48), scaled (:82-104)
    ((1,), {}, [0]), ((4,), {}, [-2, -1, 0, 1]), ((5,), {}, [-2, -1, 0, 1, 2]),
    ((4,), dict(center=2), [0, 1, 2, 3]), ((5,), dict(center=2), [0, 1, 2, 3, 4]),
    ((1, 0.5), {}, [0]), ((4, 0.5), {}, [-1, -0.5, 0, 0.5]), ((5, 0.5), {}, [-1, -0.5, 0, 0.5, 1]),
    ((4, 0.5), dict(center=1.0), [0, 0.5, 1, 1.5]), ((5, 0.5), dict(center=1.0), [0, 0.5, 1, 1.5, 2]),
    ((1, 2), {}, [0]), ((4, 2), {}, [-4, -2, 0, 2]), ((5, 2), {}, [-4, -2, 0, 2, 4]),
    ((4, 2), dict(center=4), [0, 2, 4, 6]), ((5, 2), dict(center=4), [0, 2, 4, 6, 8]),
    # first, last (not) included (:106-176)
    ((1,), dict(first=0), [0]), ((4,), dict(first=-2), [-2, -1, 0, 1]), ((5,), dict(first=-2), [-2, -1, 0, 1, 2]),
    ((4,), dict(first=0), [0, 1, 2, 3]), ((5,), dict(first=0), [0, 1, 2, 3, 4]),
    ((4,), dict(last=2), [-2, -1, 0, 1]), ((5,), dict(last=3), [-2, -1, 0, 1, 2]),
    ((4,), dict(last=4), [0, 1, 2, 3]), ((5,), dict(last=5), [0, 1, 2, 3, 4]), ((1,), dict(last=1), [0]),
    ((1,), dict(last=0, include_last=True), [0]), ((1,), dict(last=1, include_last=True), [1]),
    ((1,), dict(last=5, include_last=True), [5]), ((4,), dict(last=1, include_last=True), [-2, -1, 0, 1]),
    ((5,), dict(last=2, include_last=True), [-2, -1, 0, 1, 2]), ((4,), dict(last=3, include_last=True), [0, 1, 2, 3]),
    ((5,), dict(last=4, include_last=True), [0, 1, 2, 3, 4]),
    ((1,), dict(first=0, last=1), [0]), ((1,), dict(first=5, last=6), [5]), ((1,), dict(first=5, last=10), [5]),
    ((5,), dict(first=5, last=10), [5, 6, 7, 8, 9]),
    ((2,), dict(first=0, last=1, include_last=True), [0, 1]), ((2,), dict(first=0, last=10, include_last=True), [0, 10]),
    ((2,), dict(first=5, last=10, include_last=True), [5, 10]),
    ((6,), dict(first=5, last=10, include_last=True), [5, 6, 7, 8, 9, 10]),
]
# center on or between samples (:188-224): (shape, step, center or None, center_at_index, expected)
CENTERING = [
    (5, None, 0, True, [-2, -1, 0, 1, 2]), (5, None, 0, False, [-2, -1, 0, 1, 2]),
    (4, None, 0, True, [-2, -1, 0, 1]), (4, None, 0, False, [-1.5, -0.5, 0.5, 1.5]),
    (5, 2, 0, True, [-4, -2, 0, 2, 4]), (5, 2, 0, False, [-4, -2, 0, 2, 4]),
    (4, 2, 0, True, [-4, -2, 0, 2]), (4, 2, 0, False, [-3, -1, 1, 3]),
    (5, 0.5, 0, True, [-1.0, -0.5, 0.0, 0.5, 1.0]), (5, 0.5, 0, False, [-1.0, -0.5, 0.0, 0.5, 1.0]),
    (4, 0.5, 0, True, [-1.0, -0.5, 0.0, 0.5]), (4, 0.5, 0, False, [-0.75, -0.25, 0.25, 0.75]),
    (5, None, None, True, [-2, -1, 0, 1, 2]), (5, None, None, False, [-2, -1, 0, 1, 2]),
    (4, None, None, True, [-2, -1, 0, 1]), (4, None, None, False, [-1.5, -0.5, 0.5, 1.5]),
    (4, 2, None, False, [-3, -1, 1, 3]), (4, 0.5, None, False, [-0.75, -0.25, 0.25, 0.75]),
]


@pytest.mark.parametrize('args,kwargs,expected', COORDINATES)
def test_coordinates(args, kwargs, expected):
    npt.assert_array_equal(Grid(*args, **kwargs), np.asarray(expected), err_msg=f'Grid{args} {kwargs}')


@pytest.mark.parametrize('shape,step,center,at_index,expected', CENTERING)
def test_center_at_index(shape, step, center, at_index, expected):
    kw = {} if center is None else dict(center=center)
    npt.assert_array_equal(Grid(shape, step, center_at_index=at_index, **kw), expected)


@pytest.mark.parametrize('shape', [4, 5])
@pytest.mark.parametrize('first', [0, -2])
@pytest.mark.parametrize('at_index', [True, False])
def test_first_with_either_centering(shape, first, at_index):
    g = Grid(shape, first=first, center_at_index=at_index)
    npt.assert_equal(g.first, first)
    npt.assert_array_equal(g, first + np.arange(shape))


def test_inconsistent_first_and_last_raise():
    for first, last in ((0, 1), (0, 5), (5, 6), (5, 10)):
        with pytest.raises(ValueError):
            Grid(1, first=first, last=last, include_last=True)


def test_equality_and_extent():
    same = [(Grid(10, 0.5), Grid(10, 0.5, center=0)), (Grid(10, 0.5), Grid(10, 0.5, first=-2.5)),
            (Grid(10, 0.5), Grid(10, extent=5)), (Grid([10, 10], 0.5), Grid(10, [0.5, 0.5])),
            (Grid(256, 0.125), Grid(step=0.125, extent=32)), (Grid([256, 128], 0.125), Grid(step=0.125, extent=[32, 16])),
            (Grid(extent=[32, 16]), Grid(shape=[1, 1], extent=[32, 16]))]
    different = [(Grid(10, 0.5), Grid(10, -0.5)), (Grid(10, 0.5), Grid(10, 0.5, center=1)),
                 (Grid(10, 0.5), Grid(10, 0.5, first=0)), (Grid(10, 0.5), Grid(10, extent=6)),
                 (Grid(10, 2), Grid(10.0, 2.0))]  # the same coordinates, integer against floating point
    for a, b in same:
        assert a == b, (a, b)
    for a, b in different:
        assert not a == b, (a, b)


def test_arithmetic():
    assert -Grid(10, 0.5) == Grid(10, -0.5)
    assert -Grid(10, 0.5, first=0) == Grid(10, -0.5, first=0)
    assert -Grid(10, 0.5) == Grid(10, -0.5, first=2.5)
    for n in (1, 4, 5):
        npt.assert_array_equal(Grid(n, 0.5), Grid(n) / 2)
    npt.assert_array_equal(Grid(4, 0.5, center=1.0), (Grid.from_ranges(np.array([0, 1, 2, 3])) / 2)[0].ravel())
    npt.assert_array_equal(Grid(5, 0.5, center=1.0), (Grid.from_ranges(np.array([0, 1, 2, 3, 4])) / 2)[0].ravel())
    npt.assert_array_equal(Grid(4) + 3, [1, 2, 3, 4])
    npt.assert_array_equal(Grid(4) - 1, [-3, -2, -1, 0])
    npt.assert_array_equal(3 * Grid(3), [-3, 0, 3])
    g = Grid([2, 3], 1) @ Grid(4, 0.5)
    assert g.ndim == 3 and list(g.shape) == [2, 3, 4] and list(g.step) == [1, 1, 0.5]
    with pytest.raises(TypeError):
        Grid(3) * Grid(3)


def test_integer_arguments_keep_integer_coordinates():
    g = Grid(shape=(2, 3), first=(4, 5))
    npt.assert_array_equal(g[0], [[4], [5]])
    npt.assert_array_equal(g[1], [[5, 6, 7]])
    assert g[0].dtype == int and g[1].dtype == int


def test_frequency_grids():
    f5 = [0, 1 / 5, 2 / 5, -2 / 5, -1 / 5]
    npt.assert_array_equal(Grid(5).f, f5)
    npt.assert_array_equal(Grid(5).f.as_origin_at_0, f5)
    npt.assert_array_equal(Grid(5).f.as_origin_at_center, [-2 / 5, -1 / 5, 0, 1 / 5, 2 / 5])
    npt.assert_array_equal(Grid(5).k, np.array(f5) * 2 * np.pi)
    npt.assert_array_equal(Grid(5).k.as_origin_at_center, np.array([-2, -1, 0, 1, 2]) / 5 * 2 * np.pi)
    npt.assert_array_equal(Grid(5, 2).f, np.array(f5) / 2)
    npt.assert_array_equal(Grid([5, 2]).f[0], np.array(f5)[:, None])


def test_singleton_axes():
    g = Grid(shape=[3, 1], step=[1, 0], first=(1, 4))
    npt.assert_equal(g[0], [[1], [2], [3]])
    npt.assert_equal(g[1], [[4]])
    g = Grid(shape=[3, 1], step=[1 / 2, 0], first=(1, 4))
    npt.assert_equal(g[0], [[1], [1.5], [2]])
    npt.assert_equal(g[1], [[4]])


@pytest.mark.parametrize('ranges', [([1, 2, 3], 4), ([1, 2, 3], (4, 5)), (np.array([1, 2, 3]), np.array([4, 5])),
                                    ([[1], [2], [3]], (4, 5)), ([[1], [2], [3]], [[4, 5]]),
                                    ([[1, 1], [2, 2], [3, 3]], [[4, 5], [4, 5], [4, 5]])])
def test_from_ranges_axes(ranges):
    g = Grid.from_ranges(*ranges)
    npt.assert_equal(g[0], [[1], [2], [3]])
    npt.assert_equal(g[1], [[4, 5]] if np.size(ranges[1]) > 1 else [[4]])


@pytest.mark.parametrize('values,origin_at_center,first,extent', [
    ([-2 / 3, -1 / 3, 0, 1 / 3], True, -2 / 3, 4 / 3), ([0, 1 / 3, -2 / 3, -1 / 3], False, -2 / 3, 4 / 3),
    ([-1 / 3, 0, 1 / 3], True, -1 / 3, 1), ([0, 1 / 3, -1 / 3], False, -1 / 3, 1)])
def test_from_ranges_shifted(values, origin_at_center, first, extent):
    g = Grid.from_ranges(values)
    npt.assert_equal(g.origin_at_center, origin_at_center)
    npt.assert_equal(g.step, 1 / 3)
    npt.assert_equal(g.first, first)
    npt.assert_equal(g.center, 0)
    npt.assert_equal(g.extent, extent)
    npt.assert_equal(g.shape, len(values))
    npt.assert_equal(g[0], values)


def test_conversions_and_projections():
    g = Grid([4, 6, 8], [1, 2, 3], center=[0, 1, 2])
    assert g.as_flat[1].shape == (6,) and g.as_non_flat[1].shape == (1, 6, 1)
    npt.assert_array_equal(g.as_origin_at_0[0].ravel(), np.fft.ifftshift(g[0].ravel()))
    p = g.project([2, 0])
    assert list(p.shape) == [8, 4] and list(p.step) == [3, 1] and list(p.center) == [2, 0]
    assert list(g.transpose().shape) == [8, 6, 4]
    assert list(g.swapaxes([0, 1]).shape) == [6, 4, 8]
    assert list(g.project(axes_to_remove=1).shape) == [4, 8]
    assert g.immutable == g and g.mutable.immutable == g
    assert len(g) == 3 and len(Grid(7)) == 7 and list(Grid(3)) == [-1, 0, 1]


def test_immutable_grid_rejects_assignment():
    g = Grid(4, 1)
    for name, value in (('shape', 10), ('step', 2), ('center', 5), ('first', 0)):
        with pytest.raises(AttributeError):
            setattr(g, name, value)


def test_mutable_grid():
    g = MutableGrid(4, 1)
    g.step = 2
    assert g == Grid(4, 2).mutable
    g.step = 3.14
    assert g == Grid(4, 3.14).mutable
    g.shape = 10
    g.step = 2
    g.dtype = int
    assert g == Grid(10, 2).mutable
    g.first = 0
    assert g == Grid(10, 2, first=0).mutable
    g.center = 0
    assert g == Grid(10, 2, center=0).mutable
