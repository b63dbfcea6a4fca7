"""GPU parity of LocalDensity and PMFTXY (the other consumers of the query
path): the HIP kernels through the C-ABI against the oracle on the same
NeighborLists, plus the reference's known answers (tests/test_density_local_density.py,
tests/test_pmft.py of freud 3.5.0)."""

import numpy as np
import numpy.testing as npt
import pytest

import freud_amd
from freud_amd.box import Box
from freud_amd.pmft import PMFTXY

pytestmark = pytest.mark.gpu


def _bt(box):
    return (box.Lx, box.Ly, box.Lz, box.xy, box.xz, box.yz, box.is2D)


def test_local_density_known_answer():
    # test_density_local_density.py:80-98 of the reference
    box = Box.cube(10)
    points = np.array([[0, 0, 0], [1, 0, 0]], np.float32)
    query_points = np.array([[0, 1, 0], [-1, -1, 0], [4, 0, 0]], np.float32)
    ld = freud_amd.density.LocalDensity(2, 1)
    with pytest.raises(AttributeError):
        ld.density
    ld.compute((box, points), query_points)
    v = 4 / 3 * 8 * np.pi
    d1 = np.linalg.norm(points[1] - query_points[1])
    frac = 1 if d1 < 1.5 else (0 if d1 > 2.5 else -d1 + 2 + 0.5)
    npt.assert_allclose(ld.density, [2 / v, (1 + frac) / v, 0], rtol=1e-4)
    assert ld.box == box
    with pytest.raises(ValueError):
        freud_amd.density.LocalDensity(-1, 10)
    with pytest.raises(ValueError):
        freud_amd.density.LocalDensity(2, -1)


@pytest.mark.parametrize("is2D", [False, True])
def test_local_density_matches_oracle(oracle, is2D):
    box, pos = freud_amd.data.make_random_system(10, 10000, is2D=is2D, seed=123)
    pos = pos.astype(np.float32)
    r_max, diameter = (3.0, 1.0) if not is2D else (2.0, 0.5)
    ld = freud_amd.density.LocalDensity(r_max, diameter).compute((box, pos))
    ref_nl = oracle.query(_bt(box), pos, pos, dict(r_max=r_max + 0.5 * diameter, exclude_ii=True), nthreads=8)
    dens, num = oracle.local_density(ref_nl, len(pos), r_max, diameter, is2D)
    npt.assert_array_equal(ld.num_neighbors, num)
    npt.assert_array_equal(ld.density, dens)
    if not is2D:
        npt.assert_array_less(np.abs(ld.density - 10.0), 1.5)
    # a given NeighborList gives the same sums
    nl = freud_amd.AABBQuery(box, pos).query(pos, dict(r_max=r_max + 0.5 * diameter, exclude_ii=True)).toNeighborList()
    ld2 = freud_amd.density.LocalDensity(r_max, diameter).compute((box, pos), neighbors=nl)
    npt.assert_array_equal(ld2.density, dens)


def test_pmftxy_matches_oracle(oracle):
    box, pts = freud_amd.data.make_random_system(10, 2500, is2D=True, seed=3)
    pts = pts.astype(np.float32)
    rng = np.random.default_rng(5)
    angles = (rng.random(len(pts)) * 2 * np.pi).astype(np.float32)
    lim = np.float32(np.sqrt(9 / 2))
    pm = PMFTXY(lim, lim, (100, 90))
    pm.compute((box, pts), angles)
    ref_nl = oracle.query(_bt(box), pts, pts, dict(r_max=float(pm.r_max), exclude_ii=True), nthreads=8)
    counts = oracle.pmftxy_accumulate(ref_nl, angles, lim, lim, 100, 90)
    npt.assert_array_equal(pm.bin_counts, counts)
    pcf = oracle.pmftxy_reduce(_bt(box), counts, lim, lim, len(pts), len(pts), 1)
    npt.assert_array_equal(pm._pcf, pcf)
    with np.errstate(divide="ignore"):
        npt.assert_array_equal(pm.pmft, -np.log(pcf))
    # reset=False accumulates a second frame; the normalisation counts frames
    pm.compute((box, pts), angles, reset=False)
    npt.assert_array_equal(pm.bin_counts, 2 * counts)
    npt.assert_array_equal(pm._pcf, oracle.pmftxy_reduce(_bt(box), 2 * counts, lim, lim, len(pts), len(pts), 2))
    # bin edges and centres (test_pmft.py:385-392)
    npt.assert_allclose(pm.bin_edges[0], np.linspace(-lim, lim, 101, dtype=np.float32), atol=1e-5)
    npt.assert_allclose(pm.bin_edges[1], np.linspace(-lim, lim, 91, dtype=np.float32), atol=1e-5)
    assert pm.nbins == [100, 90]


def test_pmftxy_known_answers():
    # test_pmft.py:455-485: every query point finds the origin as its nearest neighbour
    box = Box.square(8)
    points = np.array([[0, 0, 0]], np.float32)
    query_points = np.array([[1.1, 0.0, 0.0], [-1.2, 0.0, 0.0], [0.0, 1.3, 0.0], [0.0, -1.4, 0.0]], np.float32)
    pm = PMFTXY(3, 3, 3)
    pm.compute((box, points), np.zeros(4, np.float32), query_points,
               neighbors={"mode": "nearest", "num_neighbors": 1})
    npt.assert_array_equal(pm.bin_counts, [[0, 1, 0], [1, 0, 1], [0, 1, 0]])
    pm.compute((box, query_points), np.zeros(1, np.float32), points,
               neighbors={"mode": "nearest", "num_neighbors": 1})
    npt.assert_array_equal(pm.bin_counts, [[0, 0, 0], [0, 0, 0], [0, 1, 0]])
    # test_pmft.py:487-521: the query point's orientation applies, as quaternions too
    p1 = np.array([[0.1, 0.1, 0]], np.float32)
    p2 = np.array([[1, 0, 0]], np.float32)
    pm = PMFTXY(3, 3, 12)
    quat = np.array([[1, 0, 0, 0]], np.float32)
    pm.compute((box, p1), quat, p2, neighbors={"mode": "nearest", "num_neighbors": 1})
    assert tuple(np.asarray(np.where(pm.bin_counts)).squeeze()) == (4, 6)
    pm.compute((box, p2), np.zeros(1, np.float32), p1, neighbors={"mode": "nearest", "num_neighbors": 1})
    assert tuple(np.asarray(np.where(pm.bin_counts)).squeeze()) == (7, 5)
    with pytest.raises(ValueError, match="only supports"):
        pm.compute((Box.cube(8), p1), np.zeros(1, np.float32))
