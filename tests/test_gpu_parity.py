"""GPU parity: the HIP path (through the C-ABI) against the CPU oracle on the
same seeded inputs.  Integer and index outputs and the bond vectors/distances
must be bit-identical; floating-point order parameters match within the
tolerances stated in each test."""

import os

import numpy as np
import numpy.testing as npt
import pytest

import freud_amd
from conftest import GOLDEN
from freud_amd.box import Box
from freud_amd.gsd import open_gsd

pytestmark = pytest.mark.gpu


def _bt(box):
    return (box.Lx, box.Ly, box.Lz, box.xy, box.xz, box.yz, box.is2D)


def _systems():
    rng = np.random.default_rng(11)
    out = []
    b = Box.cube(16.0)
    out.append(("cubic_uniform", b, ((rng.random((4096, 3)) - 0.5) * 16).astype(np.float32)))
    b = Box(12.0, 13.0, 14.0, 0.3, -0.2, 0.25)
    out.append(("triclinic", b, b.make_absolute(rng.random((3000, 3)).astype(np.float32))))
    b = Box.square(20.0)
    p = b.make_absolute(rng.random((2000, 3)).astype(np.float32))
    p[:, 2] = 0
    out.append(("2d", b, p))
    bx, pf = freud_amd.data.fcc_jittered(6, 1.5874, 0.05 * 1.5874, seed=1)
    out.append(("fcc", bx, pf))
    return out


SYSTEMS = _systems()


def _assert_nlist_equal(nl, ref):
    npt.assert_array_equal(nl[:], ref["ij"])
    npt.assert_array_equal(nl.distances, ref["distances"])
    npt.assert_array_equal(nl.vectors, ref["vectors"])
    npt.assert_array_equal(nl.weights, 1.0)


@pytest.mark.parametrize("name,box,pts", SYSTEMS, ids=[s[0] for s in SYSTEMS])
@pytest.mark.parametrize("kind", ["aabb", "linkcell"])
def test_ball_query_bit_exact(oracle, name, box, pts, kind, r_max=1.9):
    if kind == "aabb":
        nq = freud_amd.AABBQuery(box, pts)
    else:
        nq = freud_amd.LinkCell(box, pts, cell_width=r_max)
    for qargs in (dict(r_max=r_max, exclude_ii=True), dict(r_max=r_max, r_min=0.8), dict(r_max=1.1)):
        nl = nq.query(pts, qargs).toNeighborList()
        ref = oracle.query(_bt(box), pts, pts, qargs, nq_type=kind, cell_width=r_max, nthreads=8)
        _assert_nlist_equal(nl, ref)
        counts = np.bincount(ref["ij"][:, 0], minlength=len(pts)) if len(ref["ij"]) else np.zeros(len(pts))
        npt.assert_array_equal(nl.neighbor_counts, counts)
        seg = np.concatenate([[0], np.cumsum(counts)[:-1]])
        npt.assert_array_equal(nl.segments, np.where(counts > 0, seg, 0))


def test_ball_query_distinct_query_points(oracle):
    rng = np.random.default_rng(2)
    box = Box(10.0, 11.0, 12.0, 0.1, 0.2, -0.1)
    pts = box.make_absolute(rng.random((1500, 3)).astype(np.float32))
    qp = box.make_absolute(rng.random((700, 3)).astype(np.float32))
    for cls, kind in ((freud_amd.AABBQuery, "aabb"), (freud_amd.LinkCell, "linkcell")):
        nq = cls(box, pts) if kind == "aabb" else cls(box, pts, 2.0)
        nl = nq.query(qp, dict(r_max=2.0)).toNeighborList()
        ref = oracle.query(_bt(box), pts, qp, dict(r_max=2.0), nq_type=kind, cell_width=2.0, nthreads=8)
        _assert_nlist_equal(nl, ref)


def test_ball_query_sort_by_distance(oracle):
    name, box, pts = SYSTEMS[0]
    nl = freud_amd.AABBQuery(box, pts).query(pts, dict(r_max=2.5, exclude_ii=True)).toNeighborList(True)
    ref = oracle.query(_bt(box), pts, pts, dict(r_max=2.5, exclude_ii=True), sort_by_distance=True, nthreads=8)
    _assert_nlist_equal(nl, ref)


def test_points_outside_box(oracle):
    """Points exactly on +L/2 and slightly outside the box keep the reference's
    image arithmetic (cells are assigned through the wrapped coordinate)."""
    rng = np.random.default_rng(4)
    box = Box.cube(8.0)
    pts = ((rng.random((800, 3)) - 0.5) * 8).astype(np.float32)
    pts[:50, 0] = 4.0
    pts[50:80] *= 1.01
    nl = freud_amd.AABBQuery(box, pts).query(pts, dict(r_max=1.5, exclude_ii=True)).toNeighborList()
    ref = oracle.query(_bt(box), pts, pts, dict(r_max=1.5, exclude_ii=True), nthreads=8)
    _assert_nlist_equal(nl, ref)


@pytest.mark.parametrize("name,box,pts", SYSTEMS, ids=[s[0] for s in SYSTEMS])
@pytest.mark.parametrize("k", [1, 6, 12, 20])
def test_knn(oracle, name, box, pts, k):
    nl = freud_amd.AABBQuery(box, pts).query(pts, dict(num_neighbors=k, exclude_ii=True)).toNeighborList()
    ref = oracle.query(_bt(box), pts, pts, dict(num_neighbors=k, exclude_ii=True), nthreads=8)
    # same neighbour sets (ties at the k-th distance are broken by index in both)
    _assert_nlist_equal(nl, ref)


@pytest.mark.parametrize("kind", ["aabb", "linkcell"])
def test_knn_distance_tie(oracle, kind):
    """A constructed tie at the k-th slot: two candidates with r_sq one float
    step apart and the same sqrtf distance.  The selection follows (d, j), as
    the reference's sort by distance does, so the lower index is kept."""
    import brute

    box = Box.cube(20.0)
    pts = brute.knn_tie_system(box, kind)
    nq = freud_amd.AABBQuery(box, pts) if kind == "aabb" else freud_amd.LinkCell(box, pts, cell_width=2.0)
    for k in (3, 4):
        for by_d in (False, True):
            qa = dict(num_neighbors=k, exclude_ii=True)
            nl = nq.query(pts, qa).toNeighborList(sort_by_distance=by_d)
            ref = oracle.query(_bt(box), pts, pts, qa, nq_type=kind, cell_width=2.0, sort_by_distance=by_d)
            _assert_nlist_equal(nl, ref)
    nl = nq.query(pts, dict(num_neighbors=3, exclude_ii=True)).toNeighborList()
    assert sorted(nl.point_indices[nl.query_point_indices == 0].tolist()) == [1, 3, 4]


@pytest.mark.parametrize("name,box,pts", SYSTEMS[:2], ids=[s[0] for s in SYSTEMS[:2]])
def test_knn_r_min(oracle, name, box, pts):
    """kNN with r_min > 0, including r_min equal to an existing bond length
    (a candidate exactly at r_min is kept: sqrt(r_sq) >= r_min)."""
    d = freud_amd.AABBQuery(box, pts).query(pts, dict(num_neighbors=4, exclude_ii=True)).toNeighborList().distances
    for r_min in (0.35, float(np.sort(d)[len(d) // 3])):
        qa = dict(num_neighbors=8, r_min=r_min, exclude_ii=True)
        nl = freud_amd.AABBQuery(box, pts).query(pts, qa).toNeighborList()
        ref = oracle.query(_bt(box), pts, pts, qa, nthreads=8)
        _assert_nlist_equal(nl, ref)
        assert np.all(nl.distances >= np.float32(r_min))


def test_knn_exhaustive_paths(oracle):
    """Queries whose k-th neighbour lies beyond half the smallest plane
    distance go to the exhaustive kernel beside ordinary ones (a dense clump
    plus isolated points), and k > 64 sends every query there."""
    rng = np.random.default_rng(9)
    box = Box.cube(12.0)
    clump = (rng.standard_normal((300, 3)) * 0.5).astype(np.float32)
    lone = np.array([[6.0, 6.0, 6.0], [5.5, -5.9, 5.8], [-5.7, 5.6, -5.5], [3.0, 3.0, 3.0]], np.float32)
    pts = np.concatenate([clump, lone]).astype(np.float32)
    pts = box.wrap(pts).astype(np.float32)
    for k in (8, 80):
        qa = dict(num_neighbors=k, exclude_ii=True)
        nl = freud_amd.AABBQuery(box, pts).query(pts, qa).toNeighborList()
        ref = oracle.query(_bt(box), pts, pts, qa, nthreads=8)
        _assert_nlist_equal(nl, ref)


def test_knn_large_k(oracle):
    """num_neighbors above 64 (the exhaustive kernel) on 4096 uniform points,
    k = 65, 100 and 128; and a clear error instead of minutes of work when the
    exhaustive search would cover too many point pairs (checked on the host
    before any kernel runs)."""
    box, pts = SYSTEMS[0][1], SYSTEMS[0][2]
    aq = freud_amd.AABBQuery(box, pts)
    for k in (65, 100, 128):
        qa = dict(num_neighbors=k, exclude_ii=True)
        nl = aq.query(pts, qa).toNeighborList()
        _assert_nlist_equal(nl, oracle.query(_bt(box), pts, pts, qa, nthreads=8))
    big_box, big = freud_amd.data.fcc_jittered(40, 1.5874, 0.05 * 1.5874, seed=2)  # 256,000 points
    with pytest.raises(RuntimeError, match="exhaustive search"):
        freud_amd.AABBQuery(big_box, big).query(big, dict(num_neighbors=80)).toNeighborList()


@pytest.mark.parametrize("name,box,pts", SYSTEMS[:2] + SYSTEMS[3:], ids=[s[0] for s in SYSTEMS[:2] + SYSTEMS[3:]])
def test_knn_after_ball_query(oracle, name, box, pts):
    """A kNN query on a neighbour-query object whose ball query already built a
    grid of a suitable cell size runs on that grid; the NeighborList is the
    same as from a fresh object."""
    nq = freud_amd.AABBQuery(box, pts)
    for r_ball in (1.2, 1.5, 2.0):
        nq.query(pts, dict(r_max=r_ball, exclude_ii=True)).toNeighborList()
        for k in (6, 12):
            qa = dict(num_neighbors=k, exclude_ii=True)
            ref = oracle.query(_bt(box), pts, pts, qa, nthreads=8)
            for mode in (0, 1):  # ring and staged search on the reused grid
                with _ForcedKnn(box, len(pts), k, 0, mode):
                    nl = nq.query(pts, qa).toNeighborList()
                _assert_nlist_equal(nl, ref)


def _knn_tune_key(box, n, k):
    dens = n / box.volume
    rk = np.sqrt(2.0 * k / (np.pi * dens)) if box.is2D else np.cbrt(3.0 * 2.0 * k / (4.0 * np.pi * dens))
    return int(np.floor(2.0 * np.log2(dens * rk ** 3) + 0.5)) * 128 + k


class _ForcedKnn:
    """Presets the autotuner's kNN choice for (box, n, k): `variant` for a
    fresh grid (0-3 ring search, 4-7 staged search, by cell size) and
    `reuse_mode` on a grid another query built (0 ring, 1 staged)."""

    def __init__(self, box, n, k, variant, reuse_mode):
        from freud_amd import _lib

        self._lib = _lib
        self.key = _knn_tune_key(box, n, k)
        self.preset = {(3, self.key): variant, (4, self.key): reuse_mode}

    def __enter__(self):
        self.before = self._lib.tune_table()
        self._lib.tune_load(self.preset)

    def __exit__(self, *exc):
        self._lib.tune_load({a: self.before.get(a, -1) for a in self.preset})


@pytest.mark.parametrize("is2D", [False, True])
def test_knn_staged_paths(oracle, is2D):
    """The staged kNN search's rarer branches beside its common one: dense
    clumps (far above the mean density, so a lane's stage fills and is reduced
    to its k smallest, repeatedly), sparse regions (fewer than k points inside
    the stage radius: the point is handed back to the ring search), a radius
    limit below the stage radius (fewer than k neighbours), r_min, and query
    points other than the points."""
    rng = np.random.default_rng(31 + is2D)
    L = 30.0
    box = Box.square(L) if is2D else Box.cube(L)
    dim = 2 if is2D else 3
    centres = (rng.random((5, dim)) - 0.5) * L
    clumps = np.concatenate([c + rng.standard_normal((1400, dim)) * 0.8 for c in centres])
    spread = (rng.random((6000, dim)) - 0.5) * L
    pts = np.concatenate([clumps, spread])
    if is2D:
        pts = np.concatenate([pts, np.zeros((len(pts), 1))], axis=1)
    pts = box.wrap(pts.astype(np.float32)).astype(np.float32)
    if is2D:
        pts[:, 2] = 0
    qp = box.wrap(((rng.random((3000, 3)) - 0.5) * L).astype(np.float32)).astype(np.float32)
    if is2D:
        qp[:, 2] = 0
    nq = freud_amd.AABBQuery(box, pts)
    cases = [(pts, dict(num_neighbors=k, exclude_ii=True)) for k in (4, 8, 12, 16)]
    cases += [(pts, dict(num_neighbors=12, r_max=1.0, exclude_ii=True)),
              (pts, dict(num_neighbors=8, r_min=0.2, exclude_ii=True)),
              (qp, dict(num_neighbors=12)), (qp, dict(num_neighbors=6, r_max=2.0))]
    for q, qa in cases:
        refs = {by_d: oracle.query(_bt(box), pts, q, qa, sort_by_distance=by_d, nthreads=8) for by_d in (False, True)}
        # every staged cell size, then the ring search, on fresh grids and on the grid already built
        for variant, mode in ((4, 1), (5, 1), (6, 1), (7, 1), (0, 0)):
            print(f"staged knn 2D={is2D} {qa} nq={len(q)} variant {variant} reuse mode {mode}", flush=True)
            with _ForcedKnn(box, len(pts), qa["num_neighbors"], variant, mode):
                for fresh in (True, False):
                    obj = freud_amd.AABBQuery(box, pts) if fresh else nq
                    for by_d in (False, True):
                        nl = obj.query(q, qa).toNeighborList(sort_by_distance=by_d)
                        _assert_nlist_equal(nl, refs[by_d])


def test_knn_staged_revised_on_hand_backs(oracle):
    """Points irregular enough that the staged kNN search hands many queries
    back to the ring search (sparse points beside dense clumps): whatever the
    autotuner chose on the first call, its entry afterwards is a ring-search
    variant, on a fresh grid and on a reused one; results equal the oracle's
    throughout."""
    from freud_amd import _lib

    rng = np.random.default_rng(77)
    L = 30.0
    box = Box.cube(L)
    centres = (rng.random((5, 3)) - 0.5) * L
    clumps = np.concatenate([c + rng.standard_normal((1400, 3)) * 0.8 for c in centres])
    pts = box.wrap(np.concatenate([clumps, (rng.random((6000, 3)) - 0.5) * L]).astype(np.float32)).astype(np.float32)
    k = 12
    key = _knn_tune_key(box, len(pts), k)
    qa = dict(num_neighbors=k, exclude_ii=True)
    ref = oracle.query(_bt(box), pts, pts, qa, nthreads=8)
    before = _lib.tune_table()
    try:
        _lib.tune_load({(3, key): -1, (4, key): -1})
        nq = freud_amd.AABBQuery(box, pts)
        for _ in range(2):
            _assert_nlist_equal(nq.query(pts, qa).toNeighborList(), ref)
        assert _lib.tune_table()[(3, key)] < 4
        nq.query(pts, dict(r_max=1.6, exclude_ii=True)).toNeighborList()  # a grid the kNN can reuse
        for _ in range(2):
            _assert_nlist_equal(nq.query(pts, qa).toNeighborList(), ref)
        table = _lib.tune_table()
        assert table.get((4, key), 0) == 0
    finally:
        _lib.tune_load({(3, key): before.get((3, key), -1), (4, key): before.get((4, key), -1)})


def test_knn_linkcell(oracle):
    name, box, pts = SYSTEMS[1]
    nq = freud_amd.LinkCell(box, pts, cell_width=2.5)
    for k in (4, 12):
        nl = nq.query(pts, dict(num_neighbors=k, exclude_ii=True)).toNeighborList()
        ref = oracle.query(_bt(box), pts, pts, dict(num_neighbors=k, exclude_ii=True), nq_type="linkcell",
                           cell_width=2.5, nthreads=8)
        _assert_nlist_equal(nl, ref)


def test_knn_sparse_extended(oracle):
    # tests/test_locality_neighbor_query.py:253-300 of the reference (search past L/2)
    box = Box.cube(10)
    pts = np.array([[0, 0, 0], [1, 0, 0], [3, 0, 0], [2, 0, 0]], np.float32)
    for k in (3, 5):
        nl = freud_amd.AABBQuery(box, pts).query(pts, dict(num_neighbors=k, exclude_ii=True)).toNeighborList()
        ref = oracle.query(_bt(box), pts, pts, dict(num_neighbors=k, exclude_ii=True))
        _assert_nlist_equal(nl, ref)
    nl = freud_amd.AABBQuery(box, pts).query(pts, dict(num_neighbors=3, r_max=1.9, exclude_ii=True)).toNeighborList()
    assert [set(nl.point_indices[nl.query_point_indices == i]) for i in range(4)] == [{1}, {0, 3}, {3}, {1, 2}]


@pytest.mark.parametrize("name,box,pts", SYSTEMS, ids=[s[0] for s in SYSTEMS])
def test_rdf_counts_bit_exact(oracle, name, box, pts):
    r_max = 3.0 if name != "fcc" else 2.9
    rdf = freud_amd.density.RDF(bins=100, r_max=r_max).compute((box, pts))
    counts = oracle.rdf_accumulate(_bt(box), pts, pts, 100, r_max, exclude_ii=True, nthreads=8)
    npt.assert_array_equal(rdf.bin_counts, counts)
    ref_rdf, ref_nr = oracle.rdf_reduce(_bt(box), counts, 100, 0.0, r_max, len(pts), len(pts), 1)
    npt.assert_allclose(rdf.rdf, ref_rdf, rtol=1e-6)
    npt.assert_allclose(rdf.n_r, ref_nr, rtol=1e-6)
    # LinkCell formula through a LinkCell system
    lc = freud_amd.LinkCell(box, pts, cell_width=r_max)
    rdf2 = freud_amd.density.RDF(bins=100, r_max=r_max, r_min=0.5).compute(lc)
    counts2 = oracle.rdf_accumulate(_bt(box), pts, pts, 100, r_max, r_min=0.5, exclude_ii=True, nq_type="linkcell",
                                    cell_width=r_max, nthreads=8)
    npt.assert_array_equal(rdf2.bin_counts, counts2)


@pytest.mark.parametrize("L,r_max,n,layout", [
    (8.0, 3.9, 600, "uniform"),     # 4 cells per side: stencil wraps onto itself
    (7.0, 3.2, 500, "uniform"),     # triclinic (see below)
    (6.0, 2.9, 300, "clumped"),     # ragged cells, many empty
    (20.0, 4.0, 5000, "clumped"),
    (10.0, 2.0, 1, "uniform"),      # a single point
    (10.0, 4.9, 1500, "outside"),   # points outside the box: generic path
    (12.0, 5.9, 900, "2d"),
])
def test_rdf_self_cluster_cases(oracle, L, r_max, n, layout):
    rng = np.random.default_rng(int(L * 10) + n)
    box = Box(L, L * 1.1, L * 0.95, 0.2, -0.1, 0.15) if n == 500 else Box.cube(L)
    if layout == "2d":
        box = Box.square(L)
    f = rng.random((n, 3))
    if layout == "clumped":
        f = (np.floor(f * 3) / 3 + rng.random((n, 3)) * 0.05) % 1.0
    pts = box.make_absolute(f.astype(np.float32))
    if layout == "outside":
        pts = pts + np.float32(0.3 * L) * (rng.random((n, 3)) > 0.8).astype(np.float32)
    if layout == "2d":
        pts[:, 2] = 0
    bins = 37
    counts = oracle.rdf_accumulate(_bt(box), pts, pts, bins, r_max, r_min=0.25, exclude_ii=True, nthreads=8)
    for variant in _rdf_variants(box, len(pts), r_max):
        rdf = freud_amd.density.RDF(bins=bins, r_max=r_max, r_min=0.25).compute((box, pts))
        npt.assert_array_equal(rdf.bin_counts, counts, err_msg=f"cell variant {variant}")


def _rdf_variants(box, n, r_max):
    """Presets each self-RDF launch variant (cells r_max / R wide, r_max /
    (R F) along x, for (R, F) = (2, 1), (3, 1), (4, 1), (2, 2), (2, 4), (3, 2),
    (2, 8), (3, 4), (2, 16)) for this problem class in turn;
    the autotuner keys a class by the expected number of points in an r_max^3
    volume, in half-octaves."""
    from freud_amd import _lib

    key = int(np.floor(2.0 * np.log2(n / box.volume * r_max ** 3) + 0.5))
    count = _lib.tune_variants(1, key)
    assert count == 9
    before = _lib.tune_table()
    try:
        for v in range(count):
            _lib.tune_load({(1, key): v})
            yield v
    finally:
        _lib.tune_load({(1, key): before.get((1, key), -1)})


def test_rdf_lj_golden():
    # tests/integration/test_reader_integrations.py:27-40 of the reference
    traj = open_gsd(os.path.join(GOLDEN, "lj.gsd"))
    rdf = freud_amd.density.RDF(bins=100, r_max=5)
    for fr in traj:
        rdf.compute(fr, reset=False)
    gold = np.load(os.path.join(GOLDEN, "rdf.npz"))["rdf"]
    npt.assert_allclose(rdf.rdf, gold, rtol=1e-5, atol=1e-5)


def test_reader_integration_dcd(oracle):
    # tests/integration/test_reader_integrations.py:33-40,50-53 of the reference on
    # lj.dcd: RDF over all frames (reset=False) equals the golden, and
    # Steinhardt(6) over the 6 nearest neighbours runs on every frame
    from freud_amd.dcd import open_dcd

    traj = open_dcd(os.path.join(GOLDEN, "lj.dcd"))
    rdf = freud_amd.density.RDF(bins=100, r_max=5)
    ql = freud_amd.order.Steinhardt(6)
    for fr in traj:
        rdf.compute(fr, reset=False)
        ql.compute(fr, neighbors={"num_neighbors": 6})
    npt.assert_allclose(rdf.rdf, np.load(os.path.join(GOLDEN, "rdf.npz"))["rdf"], rtol=1e-5, atol=1e-5)
    b = fr.box
    bt = (b.Lx, b.Ly, b.Lz, b.xy, b.xz, b.yz, b.is2D)
    ref_nl = oracle.query(bt, fr.positions, fr.positions, dict(num_neighbors=6, exclude_ii=True), nthreads=8)
    ref = oracle.steinhardt(bt, fr.positions, ref_nl, [6], nthreads=8)
    npt.assert_allclose(ql.particle_order, ref["particle_order"][:, 0], atol=1e-5)


def test_rdf_nlist_and_query_points(oracle):
    rng = np.random.default_rng(7)
    box = Box.cube(12.0)
    pts = ((rng.random((2000, 3)) - 0.5) * 12).astype(np.float32)
    qp = ((rng.random((500, 3)) - 0.5) * 12).astype(np.float32)
    aq = freud_amd.AABBQuery(box, pts)
    rdf = freud_amd.density.RDF(50, 3.0).compute(aq, query_points=qp)
    counts = oracle.rdf_accumulate(_bt(box), pts, qp, 50, 3.0, exclude_ii=False, nthreads=8)
    npt.assert_array_equal(rdf.bin_counts, counts)
    nl = aq.query(qp, dict(r_max=3.0)).toNeighborList()
    rdf2 = freud_amd.density.RDF(50, 3.0).compute(aq, query_points=qp, neighbors=nl)
    npt.assert_array_equal(rdf2.bin_counts, counts)
    # kNN neighbours are binned from the NeighborList
    rdf3 = freud_amd.density.RDF(50, 3.0).compute(aq, query_points=qp, neighbors=dict(num_neighbors=8))
    ref = oracle.query(_bt(box), pts, qp, dict(num_neighbors=8), nthreads=8)
    npt.assert_array_equal(rdf3.bin_counts, oracle.rdf_bin(ref["distances"], 50, 0.0, 3.0))


# --------------------------------------------------------------------- Steinhardt
@pytest.mark.parametrize("name,box,pts", SYSTEMS[:2] + SYSTEMS[3:], ids=[s[0] for s in SYSTEMS[:2] + SYSTEMS[3:]])
def test_steinhardt(oracle, name, box, pts):
    nl = freud_amd.AABBQuery(box, pts).query(pts, dict(r_max=1.6, exclude_ii=True)).toNeighborList()
    ref_nl = oracle.query(_bt(box), pts, pts, dict(r_max=1.6, exclude_ii=True), nthreads=8)
    ls = [4, 6, 8, 10, 12]
    for kw in (dict(), dict(average=True), dict(wl=True), dict(wl=True, wl_normalize=True),
               dict(average=True, wl=True, wl_normalize=True)):
        st = freud_amd.order.Steinhardt(ls, **kw).compute((box, pts), neighbors=nl)
        ref = oracle.steinhardt(_bt(box), pts, ref_nl, ls, nthreads=8, **kw)
        ref64 = oracle.steinhardt(_bt(box), pts, ref_nl, ls, nthreads=8, f64=True, **kw)
        ok = np.isfinite(ref64["particle_order"])
        po = st.particle_order
        # tolerance 1e-5 on q_l / w_l; against the FP64 re-evaluation too
        npt.assert_allclose(po[ok], ref["particle_order"][ok], atol=1e-5, rtol=0)
        npt.assert_allclose(po[ok], ref64["particle_order"][ok], atol=1e-5, rtol=0)
        npt.assert_allclose(st.ql, ref["ql"], atol=1e-5)
        npt.assert_allclose(st.order, ref["order"], atol=1e-5)
        for a, b in zip(st.particle_harmonics, ref["qlm"]):
            npt.assert_allclose(a, b, atol=1e-5)


def test_steinhardt_low_and_high_l(oracle):
    """l = 0..3 and 14..20 (the system order reduces up to 2l + 1 = 41 values
    per particle), plain and averaged, against the oracle and its FP64 run."""
    name, box, pts = SYSTEMS[0]
    nl = freud_amd.AABBQuery(box, pts).query(pts, dict(r_max=1.6, exclude_ii=True)).toNeighborList()
    ref_nl = oracle.query(_bt(box), pts, pts, dict(r_max=1.6, exclude_ii=True), nthreads=8)
    ls = [0, 1, 2, 3, 14, 16, 20]
    for kw in (dict(), dict(average=True)):
        st = freud_amd.order.Steinhardt(ls, **kw).compute((box, pts), neighbors=nl)
        ref64 = oracle.steinhardt(_bt(box), pts, ref_nl, ls, nthreads=8, f64=True, **kw)
        ok = np.isfinite(ref64["particle_order"])
        npt.assert_allclose(st.particle_order[ok], ref64["particle_order"][ok], atol=1e-5, rtol=0)
        npt.assert_allclose(st.order, ref64["order"], atol=1e-5)


def test_steinhardt_weighted_and_fused_l(oracle):
    """Weighted bonds (random weights through NeighborList.from_arrays) against
    the oracle.  A list of distinct even l <= 12 is evaluated in one fused bond
    pass; lists with an odd or repeated l go l by l.  The fused pass performs
    the same float operations as the single-l pass, so each of its columns
    equals the single-l result bit for bit."""
    name, box, pts = SYSTEMS[1]
    ref_nl = oracle.query(_bt(box), pts, pts, dict(r_max=1.6, exclude_ii=True), nthreads=8)
    w = np.random.default_rng(5).uniform(0.2, 2.0, len(ref_nl["ij"])).astype(np.float32)
    nl = freud_amd.NeighborList.from_arrays(len(pts), len(pts), ref_nl["ij"][:, 0], ref_nl["ij"][:, 1],
                                            ref_nl["vectors"], w)
    ref_w = dict(ref_nl, weights=w)
    for ls in ([4, 6, 8, 10, 12], [12, 2, 6], [6, 6, 4], [4, 5, 6]):
        for kw in (dict(weighted=True), dict(weighted=True, average=True), dict(weighted=True, wl=True)):
            st = freud_amd.order.Steinhardt(ls, **kw).compute((box, pts), neighbors=nl)
            ref = oracle.steinhardt(_bt(box), pts, ref_w, ls, nthreads=8, **kw)
            ok = np.isfinite(ref["particle_order"])
            npt.assert_allclose(st.particle_order[ok], ref["particle_order"][ok], atol=1e-5, rtol=0)
            npt.assert_allclose(st.order, ref["order"], atol=1e-5)
            for k, l in enumerate(ls):
                one = freud_amd.order.Steinhardt(l, **kw).compute((box, pts), neighbors=nl)
                npt.assert_array_equal(st.particle_order[:, k], one.particle_order)
                npt.assert_array_equal(st.particle_harmonics[k], one.particle_harmonics)
                assert st.order[k] == one.order


def test_steinhardt_particles_without_bonds(oracle):
    """Particles without neighbours: q_lm = 0 / 0 (Steinhardt.cc:208) is NaN,
    so are their q_l and the system order that sums over all particles; the
    others match the oracle.  Single-l and fused passes, plain and averaged."""
    rng = np.random.default_rng(21)
    box = Box.cube(20.0)
    pts = ((rng.random((600, 3)) - 0.5) * 20).astype(np.float32)
    nl = freud_amd.AABBQuery(box, pts).query(pts, dict(r_max=1.2, exclude_ii=True)).toNeighborList()
    ref_nl = oracle.query(_bt(box), pts, pts, dict(r_max=1.2, exclude_ii=True), nthreads=8)
    lonely = np.bincount(ref_nl["ij"][:, 0], minlength=len(pts)) == 0
    assert lonely.any() and not lonely.all()
    for ls in ([6], [4, 6, 8]):
        for kw in (dict(), dict(average=True)):
            st = freud_amd.order.Steinhardt(ls, **kw).compute((box, pts), neighbors=nl)
            ref = oracle.steinhardt(_bt(box), pts, ref_nl, ls, nthreads=8, **kw)
            po = np.asarray(st.particle_order).reshape(len(pts), -1)
            rpo = np.asarray(ref["particle_order"]).reshape(len(pts), -1)
            npt.assert_array_equal(np.isnan(po), np.isnan(rpo))
            assert np.isnan(po[lonely]).all()
            npt.assert_allclose(po[~np.isnan(rpo)], rpo[~np.isnan(rpo)], atol=1e-5, rtol=0)
            assert np.isnan(np.ravel(st.order)).all() and np.isnan(np.ravel(ref["order"])).all()


def test_steinhardt_gc_golden():
    # tests/validation/test_steinhardt_average.py:82-109 of the reference
    fr = open_gsd(os.path.join(GOLDEN, "steinhardt_average", "Test_Configuration.gsd"))[-1]
    aq = freud_amd.AABBQuery(fr.freud_box(), fr.positions)
    nl = aq.query(fr.positions, dict(mode="ball", r_max=1.4, exclude_ii=True)).toNeighborList()
    for average, name in ((False, "GC_rc1.4_q4q6w4w6"), (True, "GC_rc1.4_avq4avq6avw4avw6")):
        gc = np.genfromtxt(os.path.join(GOLDEN, "steinhardt_average", f"{name}.txt"))
        for col, (l, wl) in enumerate([(4, False), (6, False), (4, True), (6, True)]):
            st = freud_amd.order.Steinhardt(l, average=average, wl=wl, wl_normalize=wl).compute(aq, neighbors=nl)
            npt.assert_allclose(st.particle_order, gc[:, col], atol=2e-5)


def test_steinhardt_fcc_constants():
    box, pts = freud_amd.data.UnitCell.fcc().generate_system(4, scale=2)
    for nb in (dict(r_max=1.5), dict(num_neighbors=12)):
        st = freud_amd.order.Steinhardt(6).compute((box, pts), neighbors=nb)
        npt.assert_allclose(st.particle_order, 0.57452416, atol=1e-5)
        npt.assert_allclose(st.order, 0.57452416, atol=1e-5)
        st = freud_amd.order.Steinhardt(6, wl=True).compute((box, pts), neighbors=nb)
        npt.assert_allclose(st.particle_order, -0.00262604, atol=1e-5)


@pytest.mark.parametrize("l", [1, 2, 5, 6, 11, 12, 19])
def test_steinhardt_axis_aligned(l):
    box = Box.cube(10)
    pts = np.array([[0, 0, -1], [0, 0, 0], [0, 0, 1]], np.float32)
    st = freud_amd.order.Steinhardt(l).compute((box, pts), neighbors={"num_neighbors": 2})
    npt.assert_allclose(st.particle_order, [1, 1, 1] if l % 2 == 0 else [1, 0, 1], atol=1e-5)


# --------------------------------------------------------------------- SolidLiquid, Cluster, Hexatic
def test_solidliquid(oracle):
    box, pts, _ = freud_amd.data.two_phase(5, seed=3)
    nl = freud_amd.AABBQuery(box, pts).query(pts, dict(r_max=1.5, exclude_ii=True)).toNeighborList()
    ref_nl = oracle.query(_bt(box), pts, pts, dict(r_max=1.5, exclude_ii=True), nthreads=8)
    sl = freud_amd.order.SolidLiquid(6, 0.7, 6).compute((box, pts), neighbors=nl)
    st = oracle.steinhardt(_bt(box), pts, ref_nl, [6], nthreads=8)
    ref = oracle.solidliquid(ref_nl, st["qlm"][0], st["ql"][:, 0], 6, 0.7, 6)
    npt.assert_allclose(sl.ql_ij, ref["ql_ij"], atol=1e-5)
    near = np.abs(ref["ql_ij"] - 0.7) < 1e-5
    assert not near.any(), "a bond dot product lies within the tolerance of the threshold"
    npt.assert_array_equal(sl.num_connections, ref["num_connections"])
    npt.assert_array_equal(sl.cluster_idx, ref["cluster_idx"])
    assert len(sl.cluster_sizes) == ref["num_clusters"]
    npt.assert_array_equal(sl.cluster_sizes, np.bincount(ref["cluster_idx"]))


def test_solidliquid_fcc():
    # tests/test_order_solid_liquid.py:56-75 of the reference
    box, pts = freud_amd.data.UnitCell.fcc().generate_system(4, scale=2)
    sl = freud_amd.order.SolidLiquid(6, q_threshold=0.7, solid_threshold=6).compute((box, pts),
                                                                                    neighbors=dict(r_max=1.5))
    npt.assert_array_equal(sl.num_connections, 12)
    assert sl.largest_cluster_size == len(pts)
    npt.assert_array_equal(sl.cluster_idx, 0)


def test_cluster(oracle):
    rng = np.random.default_rng(9)
    box = Box.cube(20.0)
    pts = ((rng.random((3000, 3)) - 0.5) * 20).astype(np.float32)
    cl = freud_amd.cluster.Cluster().compute((box, pts), neighbors=dict(r_max=0.9))
    ref_nl = oracle.query(_bt(box), pts, pts, dict(r_max=0.9, exclude_ii=True), nthreads=8)
    idx, ncl = oracle.cluster(len(pts), ref_nl["ij"])
    assert cl.num_clusters == ncl
    npt.assert_array_equal(cl.cluster_idx, idx)


def test_hexatic(oracle):
    box, pts = freud_amd.data.hex2d_jittered(3000, jitter=0.08, seed=1)
    hx = freud_amd.order.Hexatic(6).compute((box, pts))
    ref_nl = oracle.query(_bt(box), pts, pts, dict(num_neighbors=6, exclude_ii=True), nthreads=8)
    ref = oracle.hexatic(_bt(box), pts, ref_nl, k=6)
    npt.assert_allclose(hx.particle_order, ref, atol=1e-5)
    with pytest.raises(ValueError):
        freud_amd.order.Hexatic(6).compute((Box.cube(10), np.zeros((4, 3), np.float32)))


# --------------------------------------------------------------------- NeighborList ops
def test_nlist_ops(oracle):
    name, box, pts = SYSTEMS[1]
    aq = freud_amd.AABBQuery(box, pts)
    nl = aq.query(pts, dict(r_max=2.0, exclude_ii=True)).toNeighborList()
    ref = oracle.query(_bt(box), pts, pts, dict(r_max=2.0, exclude_ii=True), nthreads=8)
    # filter_r keeps r_min <= d < r_max (NeighborList.cc:286-307)
    c = nl.copy().filter_r(1.5, 0.5)
    keep = (ref["distances"] >= 0.5) & (ref["distances"] < 1.5)
    npt.assert_array_equal(c[:], ref["ij"][keep])
    npt.assert_array_equal(c.distances, ref["distances"][keep])
    # filter by mask
    mask = (ref["ij"][:, 1] % 3) == 0
    c2 = nl.copy().filter(mask)
    npt.assert_array_equal(c2[:], ref["ij"][mask])
    npt.assert_array_equal(c2.vectors, ref["vectors"][mask])
    # sort by distance then back (NeighborList.cc:371-399)
    sd = nl.copy().sort(by_distance=True)
    order = np.lexsort((ref["ij"][:, 1], ref["distances"], ref["ij"][:, 0]))
    npt.assert_array_equal(sd[:], ref["ij"][order])
    npt.assert_array_equal(sd.copy().sort()[:], ref["ij"])
    # find_first_index
    for i in (0, 1, 17, len(pts) - 1):
        assert nl.find_first_index(i) == int(np.searchsorted(ref["ij"][:, 0], i))
    # from_arrays round trip and validation (NeighborList.cc:51-82)
    fa = freud_amd.NeighborList.from_arrays(len(pts), len(pts), ref["ij"][:, 0], ref["ij"][:, 1], ref["vectors"])
    npt.assert_array_equal(fa[:], ref["ij"])
    npt.assert_array_equal(fa.distances, ref["distances"])
    npt.assert_array_equal(fa.neighbor_counts, nl.neighbor_counts)
    with pytest.raises(ValueError, match="sorted"):
        freud_amd.NeighborList.from_arrays(4, 4, [1, 0], [0, 1], np.zeros((2, 3)))
    with pytest.raises(ValueError, match="num_points"):
        freud_amd.NeighborList.from_arrays(4, 4, [0, 1], [0, 9], np.zeros((2, 3)))


def test_all_pairs(oracle):
    rng = np.random.default_rng(1)
    box = Box.cube(5.0)
    pts = ((rng.random((60, 3)) - 0.5) * 5).astype(np.float32)
    qp = ((rng.random((40, 3)) - 0.5) * 5).astype(np.float32)
    for q, ex in ((None, True), (qp, False), (qp, True)):
        nl = freud_amd.NeighborList.all_pairs((box, pts), query_points=q, exclude_ii=ex)
        ref = oracle.all_pairs(_bt(box), pts, pts if q is None else q, exclude_ii=ex)
        _assert_nlist_equal(nl, ref)


def test_query_errors():
    box = Box.cube(10)
    pts = np.zeros((4, 3), np.float32)
    aq = freud_amd.AABBQuery(box, pts)
    with pytest.raises(RuntimeError, match="too large"):
        aq.query(pts, dict(r_max=5.0)).toNeighborList()
    with pytest.raises(ValueError):
        aq.query(pts, dict(r_max=1.0, r_min=2.0)).toNeighborList()
    with pytest.raises(RuntimeError):
        aq.query(pts, dict(mode="ball")).toNeighborList()
    with pytest.raises(ValueError):
        freud_amd.AABBQuery(box, np.zeros((0, 3), np.float32))
    with pytest.raises(RuntimeError, match="half the box"):
        freud_amd.LinkCell(box, pts, cell_width=6.0)
    with pytest.raises(ValueError, match="z != 0"):
        freud_amd.AABBQuery(Box.square(10), np.ones((3, 3), np.float32))
    with pytest.raises(NotImplementedError):
        freud_amd.order.Steinhardt(6).compute((box, pts))


# --------------------------------------------------------------------- full size properties
def test_large_frame_properties():
    """At the benchmark size (N = 1,048,576 FCC): ball bonds are reciprocal
    (the AABB formula gives d_ij == d_ji for pairs that do not cross the box
    edge), the RDF counts below r_max of the query equal the NeighborList
    histogram, and every particle of a jittered FCC crystal has q6 near the
    perfect value."""
    box, pts = freud_amd.data.fcc_jittered(64, 1.5874, 0.05 * 1.5874, seed=1)
    aq = freud_amd.AABBQuery(box, pts)
    nl = aq.query(pts, dict(r_max=1.5, exclude_ii=True)).toNeighborList()
    ij = nl[:].astype(np.int64)
    fwd = ij[:, 0] * len(pts) + ij[:, 1]
    bwd = ij[:, 1] * len(pts) + ij[:, 0]
    # pairs present in one direction only can only sit within rounding of r_max
    odd = np.setxor1d(fwd, bwd)
    d_of = dict(zip(fwd.tolist(), nl.distances.tolist()))
    assert all(abs(d_of.get(int(k), 1.5) - 1.5) < 1e-5 for k in odd)
    assert np.all(np.diff(ij[:, 0]) >= 0)
    rdf = freud_amd.density.RDF(60, 1.5).compute(aq, neighbors=dict(r_max=1.5))
    inv = np.float32(1) / (np.float32(1.5) / np.float32(60))
    binned = np.minimum((nl.distances * inv).astype(np.int64), 59)
    npt.assert_array_equal(rdf.bin_counts, np.bincount(binned, minlength=60))
    st = freud_amd.order.Steinhardt(6).compute(aq, neighbors=nl)
    # thermal jitter (sigma = 0.05 a) lowers q6 from the perfect 0.5745 but the
    # crystal stays clearly ordered (a liquid gives ~0.1-0.3 per particle)
    q6 = st.particle_order
    assert np.all(np.isfinite(q6)) and 0.3 < q6.mean() < 0.5745
    # q_l is rotation invariant: the same frame rotated by 90 degrees about z
    rot = np.ascontiguousarray(pts[:, [1, 0, 2]] * np.array([-1, 1, 1], np.float32))
    st2 = freud_amd.order.Steinhardt(6).compute((box, rot), neighbors=dict(r_max=1.5))
    npt.assert_allclose(np.sort(st2.particle_order), np.sort(q6), atol=1e-5)


def test_stress_config_shapes(oracle):
    """Full-size triclinic (500k, density 0.85, ball r_max=2, RDF 150 bins,
    Steinhardt l = 4..12) and 2D hexagonal (250k, Hexatic k=6) frames."""
    box, pts = freud_amd.data.triclinic_uniform(500_000, 0.85, seed=2, tilt=0.3)
    aq = freud_amd.AABBQuery(box, pts)
    nl = aq.query(pts, dict(r_max=2.0, exclude_ii=True)).toNeighborList()
    ref = oracle.query(_bt(box), pts, pts, dict(r_max=2.0, exclude_ii=True), nthreads=16)
    _assert_nlist_equal(nl, ref)
    rdf = freud_amd.density.RDF(150, 2.0).compute(aq)
    npt.assert_array_equal(rdf.bin_counts, oracle.rdf_accumulate(_bt(box), pts, pts, 150, 2.0, nthreads=16))
    ls = [4, 6, 8, 10, 12]
    st = freud_amd.order.Steinhardt(ls).compute(aq, neighbors=nl)
    r = oracle.steinhardt(_bt(box), pts, ref, ls, nthreads=16)
    npt.assert_allclose(st.particle_order, r["particle_order"], atol=1e-5)
    hbox, hpts = freud_amd.data.hex2d_jittered(250_000, jitter=0.05, seed=3)
    hx = freud_amd.order.Hexatic(6).compute((hbox, hpts))
    href = oracle.query(_bt(hbox), hpts, hpts, dict(num_neighbors=6, exclude_ii=True), nthreads=16)
    npt.assert_allclose(hx.particle_order, oracle.hexatic(_bt(hbox), hpts, href, k=6), atol=1e-5)


def test_spatially_ordered_paths(oracle):
    """Above 65536 points the row-gather kernels (averaged q_lm, SolidLiquid
    dots, the classifier's fallback) run in a cached spatial order when the
    points are not stored coherently, while the bond passes (q_lm, fused q_lm)
    stage each tile's bonds in index order; per-particle results must not
    change."""
    box, pts, _ = freud_amd.data.two_phase(27, seed=8)  # 157,464 points, liquid half in random order
    assert len(pts) >= 65536
    nl = freud_amd.AABBQuery(box, pts).query(pts, dict(r_max=1.5, exclude_ii=True)).toNeighborList()
    ref_nl = oracle.query(_bt(box), pts, pts, dict(r_max=1.5, exclude_ii=True), nthreads=16)
    sl = freud_amd.order.SolidLiquid(6, 0.7, 6).compute((box, pts), neighbors=nl)
    st = oracle.steinhardt(_bt(box), pts, ref_nl, [6], nthreads=16)
    ref = oracle.solidliquid(ref_nl, st["qlm"][0], st["ql"][:, 0], 6, 0.7, 6)
    npt.assert_allclose(sl.ql_ij, ref["ql_ij"], atol=1e-5)
    near = np.abs(ref["ql_ij"] - 0.7) < 1e-5
    ok = ~np.isin(np.arange(len(pts)), ref_nl["ij"][near, 0])
    npt.assert_array_equal(sl.num_connections[ok], ref["num_connections"][ok])
    if not near.any():
        npt.assert_array_equal(sl.cluster_idx, ref["cluster_idx"])
    for kw in (dict(average=True), dict(wl=True, wl_normalize=True)):
        s2 = freud_amd.order.Steinhardt(6, **kw).compute((box, pts), neighbors=nl)
        r2 = oracle.steinhardt(_bt(box), pts, ref_nl, [6], nthreads=16, **kw)
        npt.assert_allclose(s2.particle_order, np.ravel(r2["particle_order"]), atol=1e-5)
        npt.assert_allclose(s2.order, np.ravel(r2["order"])[0], atol=1e-5)
    # several even l: the fused bond pass in the spatial order
    s3 = freud_amd.order.Steinhardt([4, 6, 8, 12]).compute((box, pts), neighbors=nl)
    r3 = oracle.steinhardt(_bt(box), pts, ref_nl, [4, 6, 8, 12], nthreads=16)
    npt.assert_allclose(s3.particle_order, r3["particle_order"], atol=1e-5)
    npt.assert_allclose(s3.order, r3["order"], atol=1e-5)


@pytest.mark.parametrize("name,box,pts", SYSTEMS[:3], ids=[s[0] for s in SYSTEMS[:3]])
def test_knn_cell_variants(oracle, name, box, pts):
    """Every kNN variant the autotuner can pick (four cell sizes, ring or
    staged search) gives the same NeighborList."""
    from freud_amd import _lib

    for k in (6, 12):
        ref = oracle.query(_bt(box), pts, pts, dict(num_neighbors=k, exclude_ii=True), nthreads=8)
        dens = len(pts) / box.volume
        rk = np.sqrt(2.0 * k / (np.pi * dens)) if box.is2D else np.cbrt(3.0 * 2.0 * k / (4.0 * np.pi * dens))
        key = int(np.floor(2.0 * np.log2(dens * rk ** 3) + 0.5)) * 128 + k
        before = _lib.tune_table()
        try:
            for v in range(_lib.tune_variants(3, key)):
                _lib.tune_load({(3, key): v})
                nl = freud_amd.AABBQuery(box, pts).query(pts, dict(num_neighbors=k, exclude_ii=True)).toNeighborList()
                _assert_nlist_equal(nl, ref)
        finally:
            _lib.tune_load({(3, key): before.get((3, key), -1)})
