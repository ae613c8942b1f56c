"""Freezeout surface elements of one cell (vh_surface_square, vh_surface_cube)
against the reference's own Cornelius fixtures
(freezeout/cornelius-c++-1.3/output_{2,3,4}d.dat, kept gzipped in
tests/golden/cornelius/): the reference's example draws 10,000 squares, then
10,000 cubes, then 10,000 hypercubes, with corner values rand()/RAND_MAX from the C library's
generator in its default state, and writes each element's centroid and
normal with six significant digits.  The same draws here must give the same
elements in the same order.  Plus the closed forms: a linear field gives one
plane element whose normal is the cell's cross-section along the gradient."""
import ctypes
import gzip
import os

import numpy as np

import vhtest as V
import cpuvh_amd as C

GOLD = os.path.join(V.REPO, "tests", "golden", "cornelius")


def _fixture(dim):
    with gzip.open(os.path.join(GOLD, f"output_{dim}d.dat.gz"), "rt") as f:
        return [ln.rstrip("\n") for ln in f]


def _draws():
    libc = ctypes.CDLL(None)
    libc.rand.restype = ctypes.c_int
    libc.srand(1)  # the generator's default state (the example never seeds it)
    rand_max = 2147483647.0  # RAND_MAX of the C library
    return lambda n: np.array([libc.rand() / rand_max for _ in range(n)])


def _fmt(row):
    return "".join("%g " % x for x in row)


def _compare(dim, cells, draw):
    want = _fixture(dim)
    got = []
    for _ in range(cells):
        corners = draw(2 ** dim)
        for row in C.surface_elements(corners, 0.5, np.ones(dim)):
            got.append(_fmt(row))
    assert len(got) == len(want), (len(got), len(want))
    bad = [k for k, (a, b) in enumerate(zip(got, want)) if a != b]
    assert not bad, [(k, got[k], want[k]) for k in bad[:5]]


def test_cornelius_fixtures_2d_then_3d_then_4d():
    draw = _draws()
    _compare(2, 10000, draw)
    _compare(3, 10000, draw)
    _compare(4, 10000, draw)


def test_linear_field_gives_one_plane_element():
    rng = np.random.default_rng(5)
    for _ in range(50):
        g = rng.normal(size=3)
        d = rng.uniform(0.5, 2.0, 3)
        x = np.array([[[(i * d[0], j * d[1], k * d[2]) for k in range(2)] for j in range(2)] for i in range(2)])
        vals = x @ g
        level = float(vals.min() + (vals.max() - vals.min()) * rng.uniform(0.2, 0.8))
        el = C.surface_elements(vals.reshape(-1), level, d)
        assert el.shape[0] == 1
        c, n = el[0, :3], el[0, 3:]
        assert abs(c @ g - level) < 1e-12 * (1 + abs(level))          # the centroid lies on the plane
        assert np.allclose(np.cross(n, g), 0, atol=1e-12 * np.linalg.norm(g) * (1 + np.linalg.norm(n)))
        assert n @ g < 0                                               # towards the side below the level
