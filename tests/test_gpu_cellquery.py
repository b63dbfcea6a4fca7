"""GPU parity of freud.locality.CellQuery: ball queries and the fused RDF
through the C-ABI against the oracle's CellQuery restatement (host ghost grid,
27-cell walk), bit for bit, plus the reference's own CellQuery tests
(tests/test_locality_neighbor_query.py:697-1205 of freud 3.5.0)."""

import numpy as np
import numpy.testing as npt
import pytest

import freud_amd
from freud_amd.box import Box

pytestmark = pytest.mark.gpu


def _bt(box):
    return (box.Lx, box.Ly, box.Lz, box.xy, box.xz, box.yz, box.is2D)


def _assert_nlist_equal(nl, ref):
    npt.assert_array_equal(nl[:], ref["ij"])
    npt.assert_array_equal(nl.distances, ref["distances"])
    npt.assert_array_equal(nl.vectors, ref["vectors"])
    npt.assert_array_equal(nl.weights, 1.0)


def _systems():
    rng = np.random.default_rng(21)
    out = []
    b = Box.cube(16.0)
    out.append(("cubic", b, ((rng.random((4096, 3)) - 0.5) * 16).astype(np.float32)))
    b = Box(12.0, 13.0, 14.0, 0.3, -0.2, 0.25)
    out.append(("triclinic", b, b.make_absolute(rng.random((3000, 3)).astype(np.float32))))
    b = Box.square(20.0)
    p = b.make_absolute(rng.random((2000, 3)).astype(np.float32))
    p[:, 2] = 0
    out.append(("2d", b, p))
    # points on the faces and slightly outside the box
    b = Box(9.0, 9.5, 10.0, 0.1, 0.0, -0.2)
    f = rng.random((1500, 3)).astype(np.float32)
    f[:100, 0] = 0.0
    f[100:200, 1] = 1.0
    f[200:300] = np.where(rng.random((100, 3)) < 0.5, -0.003, 1.003).astype(np.float32)
    out.append(("faces_outside", b, b.make_absolute(f)))
    return out


SYSTEMS = _systems()


@pytest.mark.parametrize("name,box,pts", SYSTEMS, ids=[s[0] for s in SYSTEMS])
def test_cellquery_ball_bit_exact(oracle, name, box, pts):
    print(f"cellquery ball {name}", flush=True)
    nq = freud_amd.CellQuery(box, pts)
    for qargs in (dict(r_max=1.9, exclude_ii=True), dict(r_max=2.3, r_min=0.8), dict(r_max=1.1)):
        nl = nq.query(pts, qargs).toNeighborList()
        ref = oracle.query(_bt(box), pts, pts, qargs, nq_type="cellquery", nthreads=8)
        _assert_nlist_equal(nl, ref)
    # sorted by distance
    nl = nq.query(pts, dict(r_max=1.9, exclude_ii=True)).toNeighborList(sort_by_distance=True)
    ref = oracle.query(_bt(box), pts, pts, dict(r_max=1.9, exclude_ii=True), nq_type="cellquery",
                       sort_by_distance=True, nthreads=8)
    _assert_nlist_equal(nl, ref)
    # distinct query points, some outside the grid
    rng = np.random.default_rng(len(pts))
    qp = box.make_absolute((rng.random((700, 3)) * 1.6 - 0.3).astype(np.float32))
    if box.is2D:
        qp[:, 2] = 0
    nl = nq.query(qp, dict(r_max=2.0)).toNeighborList()
    ref = oracle.query(_bt(box), pts, qp, dict(r_max=2.0), nq_type="cellquery", nthreads=8)
    _assert_nlist_equal(nl, ref)


@pytest.mark.parametrize("name,box,pts", SYSTEMS, ids=[s[0] for s in SYSTEMS])
def test_cellquery_rdf_bit_exact(oracle, name, box, pts):
    print(f"cellquery rdf {name}", flush=True)
    nq = freud_amd.CellQuery(box, pts)
    rdf = freud_amd.density.RDF(bins=100, r_max=3.0, r_min=0.25).compute(nq)
    counts = oracle.rdf_accumulate(_bt(box), pts, pts, 100, 3.0, r_min=0.25, exclude_ii=True, nq_type="cellquery",
                                   nthreads=8)
    npt.assert_array_equal(rdf.bin_counts, counts)
    qp = pts[::3].copy()
    qp[:, :2] += np.float32(0.17)
    rdf = freud_amd.density.RDF(bins=64, r_max=2.5).compute(nq, query_points=qp)
    counts = oracle.rdf_accumulate(_bt(box), pts, qp, 64, 2.5, exclude_ii=False, nq_type="cellquery", nthreads=8)
    npt.assert_array_equal(rdf.bin_counts, counts)


@pytest.mark.parametrize("seed", range(10))
@pytest.mark.parametrize("is2D", [True, False])
def test_cellquery_equals_aabb(seed, is2D):
    # TestAABBQueryVsCellQuery (test_locality_neighbor_query.py:1082-1205 of the reference)
    rng = np.random.RandomState(seed)
    L = rng.uniform(10, 50) if is2D else rng.uniform(10, 30)
    N = rng.randint(10, 200) if is2D else rng.randint(10, 100)
    box, pts = freud_amd.data.make_random_system(L, N, is2D=is2D, seed=seed)
    r_max = L * rng.uniform(0.1, 0.4)
    _, qpts = freud_amd.data.make_random_system(L, 50, is2D=is2D, seed=seed + 10000)
    aq = freud_amd.AABBQuery(box, pts)
    cq = freud_amd.CellQuery(box, pts)
    for qp, ex in ((pts, True), (qpts, False)):
        a = aq.query(qp, dict(r_max=r_max, exclude_ii=ex)).toNeighborList()
        c = cq.query(qp, dict(r_max=r_max, exclude_ii=ex)).toNeighborList()
        npt.assert_array_equal(c[:], a[:])
        npt.assert_allclose(c.distances, a.distances, rtol=1e-5)
        npt.assert_allclose(c.vectors, a.vectors, rtol=1e-5, atol=1e-5)


def test_cellquery_reference_cases():
    box = Box.cube(10)
    # test_query_mode_inference and test_r_min
    pts = np.array([[0, 0, 0], [1, 0, 0], [3, 0, 0], [2, 0, 0]], np.float32)
    nq = freud_amd.CellQuery(box, pts)
    r1 = list(nq.query(pts, dict(mode="ball", r_max=2.01)))
    r2 = list(nq.query(pts, dict(r_max=2.01)))
    assert r1 == r2 and len(r1) == 14
    res = list(nq.query(pts, dict(mode="ball", r_max=2.9, r_min=1.1, exclude_ii=True)))
    got = [{b[1] for b in res if b[0] == i} for i in range(4)]
    assert got == [{3}, {2}, {1}, {0}]
    with pytest.raises(RuntimeError, match="CellQuery only supports"):
        list(nq.query(pts, dict(mode="nearest", num_neighbors=3)))
    with pytest.raises(RuntimeError, match="CellQuery only supports"):
        nq.query(pts, dict(num_neighbors=2, exclude_ii=True)).toNeighborList()
    # test_duplicate_cell_shells / test_duplicate_cell_shells2
    nq = freud_amd.CellQuery(Box.square(5), [[0.9, 0, 0]])
    assert len(list(nq.query([[-1.5, 0, 0]], dict(r_max=2.45)))) == 1
    p2 = np.array([[1.5132198, 6.67087, 3.1856632], [1.3913784, -2.3667011, 4.5227165],
                   [-3.6133137, 9.043476, 0.8957424]], np.float32)
    nq = freud_amd.CellQuery(Box.cube(21), p2)
    assert len(list(nq.query(p2[0], dict(r_max=10)))) == 3
    # test_invalid_r_max_r_min_bounds
    nq = freud_amd.CellQuery(Box(3, 4, 5, 1, 0.5, 0.1), np.array([[0, 0, 0], [1, 1, 0]]))
    for qa in (dict(r_max=0), dict(r_max=-0.5), dict(r_max=0.1, r_min=0.1), dict(r_max=0.1, r_min=0.2)):
        with pytest.raises(ValueError):
            list(nq.query(np.array([[0, 0, 0], [1, 1, 0]]), qa))
    # test_too_large_r_max_raises
    nq = freud_amd.CellQuery(Box.square(5), [[0, 0, 0], [1, 1, 0], [1, -1, 0]])
    with pytest.raises(RuntimeError):
        list(nq.query([[0, 0, 0], [1, 1, 0], [1, -1, 0]], dict(r_max=5)))
    # test_query_point_outside_grid_bounds_ball
    nq = freud_amd.CellQuery(Box.cube(20), np.array([[0.0, 0.0, 0.0]], np.float32))
    assert len(list(nq.query(np.array([[-15.0, 0.0, 0.0]], np.float32), dict(mode="ball", r_max=5.0)))) == 0
    # test_chaining
    box, points = freud_amd.data.make_random_system(10, 500, seed=1)
    nl1 = freud_amd.CellQuery(box, points).query(points, dict(r_max=1, exclude_ii=True)).toNeighborList()
    nl2 = freud_amd.CellQuery(box, points).query(points, dict(r_max=1, exclude_ii=True)).toNeighborList()
    npt.assert_array_equal(nl1[:], nl2[:])
    npt.assert_array_equal(nl1.vectors, nl2.vectors)


def test_cellquery_large_frame(oracle):
    """A 250,000-point frame at r_max 2.5, self and sorted by distance, against
    the oracle."""
    print("cellquery large frame", flush=True)
    box, pts = freud_amd.data.make_random_system(63.0, 250000, seed=8)
    pts = pts.astype(np.float32)
    nq = freud_amd.CellQuery(box, pts)
    nl = nq.query(pts, dict(r_max=2.5, exclude_ii=True)).toNeighborList()
    ref = oracle.query(_bt(box), pts, pts, dict(r_max=2.5, exclude_ii=True), nq_type="cellquery", nthreads=16)
    _assert_nlist_equal(nl, ref)
    rdf = freud_amd.density.RDF(bins=200, r_max=2.5).compute(nq)
    npt.assert_array_equal(rdf.bin_counts, oracle.rdf_bin(ref["distances"], 200, 0.0, 2.5))
