"""GPU parity of the NeighborList filters (fa_filter_sann / fa_filter_rad, the
device form of freud/locality/FilterSANN.cc and FilterRAD.cc) against the
oracle on the same unfiltered lists: kept bonds, their order, vectors,
distances and weights bit-identical, the unfilled-shell masks equal, and the
reference's own known answers and error behaviour
(tests/test_locality_filter.py of the reference)."""

import numpy as np
import numpy.testing as npt
import pytest

import freud_amd
from freud_amd import data
from freud_amd.box import Box
from freud_amd.locality import FilterRAD, FilterSANN, NeighborList

pytestmark = pytest.mark.gpu


def _bt(box):
    return (box.Lx, box.Ly, box.Lz, box.xy, box.xz, box.yz, box.is2D)


def _systems():
    rng = np.random.default_rng(5)
    out = []
    b = Box.cube(14.0)
    out.append(("cubic_uniform", b, ((rng.random((2744, 3)) - 0.5) * 14).astype(np.float32), 1.9))
    b = Box(12.0, 13.0, 14.0, 0.3, -0.2, 0.25)
    out.append(("triclinic", b, b.make_absolute(rng.random((2184, 3)).astype(np.float32)), 1.9))
    b = Box.square(30.0)
    p = b.make_absolute(rng.random((900, 3)).astype(np.float32))
    p[:, 2] = 0
    out.append(("2d", b, p, 2.6))
    bx, pf = data.fcc_jittered(6, 1.5874, 0.05 * 1.5874, seed=2)
    out.append(("fcc", bx, pf, 1.9))
    return out


SYSTEMS = _systems()
IDS = [s[0] for s in SYSTEMS]


def _host_nlist(nl):
    return {"ij": np.asarray(nl[:]), "distances": np.asarray(nl.distances), "weights": np.asarray(nl.weights),
            "vectors": np.asarray(nl.vectors)}


def _assert_filtered_equal(got, ref, unfilled_ref=None):
    npt.assert_array_equal(got[:], ref["ij"])
    npt.assert_array_equal(got.distances, ref["distances"])
    npt.assert_array_equal(got.vectors, ref["vectors"])
    npt.assert_array_equal(got.weights, ref["weights"])
    if len(ref["distances"]):
        assert got.num_query_points == ref["num_query_points"]
        assert got.num_points == ref["num_points"]
        counts = np.bincount(ref["ij"][:, 0], minlength=ref["num_query_points"])
        npt.assert_array_equal(got.neighbor_counts, counts)
        # NeighborList.cc:199-229: a query point without bonds keeps segment 0
        first = np.concatenate([[0], np.cumsum(counts)[:-1]])
        npt.assert_array_equal(got.segments, np.where(counts > 0, first, 0))


@pytest.mark.parametrize("name,box,pts,r_max", SYSTEMS, ids=IDS)
def test_sann_matches_oracle(oracle, name, box, pts, r_max):
    nq = freud_amd.AABBQuery(box, pts)
    filt = FilterSANN(allow_incomplete_shell=True)
    filt.compute(nq, dict(r_max=r_max, exclude_ii=True))
    unfiltered = filt.unfiltered_nlist
    ref, unfilled = oracle.filter_sann(_host_nlist(unfiltered), len(pts))
    print(f"sann {name}: {len(unfiltered)} -> {len(filt.filtered_nlist)} bonds, {int(unfilled.sum())} unfilled",
          flush=True)
    _assert_filtered_equal(filt.filtered_nlist, ref)
    # the unfiltered list is the query's, in its own (i, j) order
    npt.assert_array_equal(unfiltered[:], oracle.query(_bt(box), pts, pts, dict(mode="ball", r_max=r_max,
                                                                                  exclude_ii=True))["ij"])


@pytest.mark.parametrize("name,box,pts,r_max", SYSTEMS, ids=IDS)
@pytest.mark.parametrize("closed", [True, False], ids=["closed", "open"])
@pytest.mark.parametrize("with_radii", [False, True], ids=["unit", "radii"])
def test_rad_matches_oracle(oracle, name, box, pts, r_max, closed, with_radii):
    nq = freud_amd.AABBQuery(box, pts)
    radii = np.linspace(0.75, 1.25, len(pts)).astype(np.float32) if with_radii else None
    filt = FilterRAD(allow_incomplete_shell=True, terminate_after_blocked=closed)
    filt.compute(nq, dict(r_max=r_max, exclude_ii=True), points_radii=radii)
    ref, unfilled = oracle.filter_rad(_bt(box), pts, pts, _host_nlist(filt.unfiltered_nlist),
                                      terminate_after_blocked=closed, points_radii=radii)
    print(f"rad {name} closed={closed} radii={with_radii}: {len(filt.filtered_nlist)} bonds, "
          f"{int(unfilled.sum())} unfilled", flush=True)
    _assert_filtered_equal(filt.filtered_nlist, ref)


def test_unfilled_masks_match_oracle(oracle, capsys):
    # a short cut-off leaves most shells open: the warning lists exactly the
    # oracle's unfilled query points
    box, pts = data.make_random_system(10, 300, seed=4)
    pts = np.asarray(pts, np.float32)
    for filt, run in ((FilterSANN(True), lambda nl: oracle.filter_sann(nl, len(pts))),
                      (FilterRAD(True, True), lambda nl: oracle.filter_rad(_bt(box), pts, pts, nl))):
        filt.compute((box, pts), dict(r_max=1.3, exclude_ii=True))
        ref, unfilled = run(_host_nlist(filt.unfiltered_nlist))
        _assert_filtered_equal(filt.filtered_nlist, ref)
        out = capsys.readouterr().out
        idx = np.flatnonzero(unfilled)
        assert len(idx)
        assert out.strip() == ("WARNING: Query point indices " + ", ".join(map(str, idx))
                               + " do not have full neighbor shells.")


def test_incomplete_shell_raises():
    # tests/test_locality_filter.py:65-72 of the reference
    sys_ = data.make_random_system(10, 5, seed=1)
    for filt in (FilterSANN(), FilterRAD()):
        with pytest.raises(RuntimeError, match=r"^Query point indices .* do not have full neighbor shells\.$"):
            filt.compute(sys_, dict(r_max=1.2, exclude_ii=True))


def test_all_pairs_default_and_query_points(oracle):
    # neighbors=None filters every pair (without i == i for self queries)
    box, pts = data.make_random_system(10, 100, seed=1)
    pts = np.asarray(pts, np.float32)
    filt = FilterSANN()
    filt.compute((box, pts))
    assert len(filt.unfiltered_nlist) == 100 * 99
    ref_all = oracle.all_pairs(_bt(box), pts, pts, exclude_ii=True)
    npt.assert_array_equal(filt.unfiltered_nlist[:], ref_all["ij"])
    ref, _ = oracle.filter_sann(ref_all, len(pts))
    _assert_filtered_equal(filt.filtered_nlist, ref)
    # separate query points: RAD reads them for wrap(q_i - p_j)
    rng = np.random.default_rng(9)
    q = box.make_absolute(rng.random((40, 3)).astype(np.float32))
    for closed in (True, False):
        filt = FilterRAD(terminate_after_blocked=closed)
        filt.compute((box, pts), query_points=q)
        ref_q = oracle.all_pairs(_bt(box), pts, q, exclude_ii=False)
        assert len(filt.unfiltered_nlist) == 40 * 100
        ref, _ = oracle.filter_rad(_bt(box), pts, q, ref_q, terminate_after_blocked=closed)
        _assert_filtered_equal(filt.filtered_nlist, ref)


def test_knn_and_given_list(oracle):
    # a kNN list and a NeighborList passed in directly, with weights SANN keeps
    box, pts = data.make_random_system(12, 500, seed=7)
    pts = np.asarray(pts, np.float32)
    nq = freud_amd.AABBQuery(box, pts)
    nl = nq.query(pts, dict(num_neighbors=24, exclude_ii=True)).toNeighborList()
    w = np.random.default_rng(1).random(len(nl)).astype(np.float32) + 0.5
    given = NeighborList.from_arrays(len(pts), len(pts), nl.query_point_indices, nl.point_indices, nl.vectors, w)
    filt = FilterSANN(True).compute(nq, given)
    assert filt.unfiltered_nlist is given
    ref, _ = oracle.filter_sann(_host_nlist(given), len(pts))
    _assert_filtered_equal(filt.filtered_nlist, ref)
    assert not np.all(filt.filtered_nlist.weights == 1)
    filt = FilterRAD(True, False).compute(nq, given)
    ref, _ = oracle.filter_rad(_bt(box), pts, pts, _host_nlist(given), terminate_after_blocked=False)
    _assert_filtered_equal(filt.filtered_nlist, ref)


def test_rad_simple_known_answer():
    # tests/test_locality_filter.py:200-257 of the reference
    pts = np.array([[0, 0, 0], [0, 2, 0], [0, 0, 4], [0, 1, 0], [0, 0, 2.1]], np.float32)
    sys_ = (Box.cube(10), pts)
    nb = dict(r_max=2.5, exclude_ii=True)
    sol = FilterRAD(allow_incomplete_shell=True, terminate_after_blocked=True).compute(sys_, nb).filtered_nlist
    npt.assert_allclose(sol.distances, [1, 1, 1.9, 1, 1, 1.9, 2.1], rtol=1e-6)
    npt.assert_array_equal(sol.point_indices, [3, 3, 4, 0, 1, 2, 0])
    npt.assert_array_equal(sol.query_point_indices, [0, 1, 2, 3, 3, 4, 4])
    sol = FilterRAD(allow_incomplete_shell=True, terminate_after_blocked=False).compute(sys_, nb).filtered_nlist
    npt.assert_allclose(sol.distances, [1, 2.1, 1, 1.9, 1, 1, 1.9, 2.1], rtol=1e-6)
    npt.assert_array_equal(sol.point_indices, [3, 4, 3, 4, 0, 1, 2, 0])
    npt.assert_array_equal(sol.query_point_indices, [0, 0, 1, 2, 3, 3, 4, 4])


def test_sann_simple_known_answer():
    # tests/test_locality_filter.py:377-447 of the reference
    pts = np.array([[0, 0, 0], [1, 0, 0], [0, 2, 0], [0, 0, 4], [0, 1, 0], [0, 0, 1]], np.float32)
    f = FilterSANN(allow_incomplete_shell=True).compute((Box.cube(10), pts), dict(r_max=1.5, exclude_ii=True))
    s2 = np.sqrt(2)
    sol = f.filtered_nlist
    npt.assert_allclose(sol.distances, [1, 1, 1, 1, s2, s2, 1, 1, 1, s2, s2, 1, s2, s2], rtol=1e-6)
    npt.assert_array_equal(sol.point_indices, [1, 4, 5, 0, 4, 5, 4, 0, 2, 1, 5, 0, 1, 4])
    npt.assert_array_equal(sol.query_point_indices, [0, 0, 0, 1, 1, 1, 2, 4, 4, 4, 4, 5, 5, 5])
    sol = f.unfiltered_nlist
    npt.assert_array_equal(sol.point_indices, [1, 4, 5, 0, 4, 5, 4, 0, 1, 2, 5, 0, 1, 4])
    npt.assert_array_equal(sol.query_point_indices, [0, 0, 0, 1, 1, 1, 2, 4, 4, 4, 4, 5, 5, 5])


@pytest.mark.parametrize(("cls", "expected"), [("bcc", 14), ("fcc", 12)])
@pytest.mark.parametrize("filt", ["sann", "rad_closed", "rad_open"])
def test_known_crystals(cls, expected, filt):
    # tests/test_locality_filter.py:90-106 of the reference
    sys_ = getattr(data.UnitCell, cls)().generate_system(3)
    f = FilterSANN() if filt == "sann" else FilterRAD(terminate_after_blocked=filt == "rad_closed")
    f.compute(sys_, neighbors=dict(r_max=1.49, exclude_ii=True))
    npt.assert_array_equal(f.filtered_nlist.neighbor_counts, expected)


def test_radii_validation():
    # tests/test_locality_filter.py:289-327 of the reference
    n = 20
    sys_ = data.make_random_system(10, n, seed=2)
    filt = FilterRAD()
    with pytest.raises(ValueError, match=rf"^array\.shape= \({n - 1},\); expected shape = \({n}\)$"):
        filt.compute(sys_, dict(r_max=4.0, exclude_ii=True), points_radii=np.ones(n - 1))
    bad = np.ones(n)
    bad[0] = 0
    with pytest.raises(ValueError, match=r"^points_radii must contain only values greater than 0\.$"):
        filt.compute(sys_, dict(r_max=4.0, exclude_ii=True), points_radii=bad)
    bad = np.ones(n)
    bad[1] = np.nan
    with pytest.raises(ValueError, match=r"^query_points_radii must contain only finite values\.$"):
        filt.compute(sys_, dict(r_max=4.0, exclude_ii=True), query_points_radii=bad)
    with pytest.raises(ValueError, match=r"^points_radii must contain only finite values\.$"):
        filt.compute(sys_, dict(r_max=4.0, exclude_ii=True), points_radii=np.full(n, np.inf))


def test_large_frame(oracle):
    # 108,000-point jittered FCC frame, ball r 1.6 (about 18 candidates each)
    box, pts = data.fcc_jittered(30, 1.5874, 0.05 * 1.5874, seed=3)
    nq = freud_amd.AABBQuery(box, pts)
    nl = nq.query(pts, dict(r_max=1.6, exclude_ii=True)).toNeighborList()
    host = _host_nlist(nl)
    print(f"large frame: {len(pts)} points, {len(nl)} bonds", flush=True)
    f = FilterSANN(True).compute(nq, nl)
    ref, unfilled = oracle.filter_sann(host, len(pts))
    _assert_filtered_equal(f.filtered_nlist, ref)
    for closed in (True, False):
        f = FilterRAD(True, closed).compute(nq, nl)
        ref, _ = oracle.filter_rad(_bt(box), pts, pts, host, terminate_after_blocked=closed)
        _assert_filtered_equal(f.filtered_nlist, ref)


@pytest.mark.parametrize("r_max", [2.2, 2.9, 3.6, 4.6, 5.5])
def test_segment_widths(oracle, r_max):
    # 3,000 random points in a 16^3 box: the longest candidate lists are 54,
    # 105, 179, 344 and 570 bonds, one case per LDS width of the per-segment
    # path (64, 128, 256, 512 bonds) and one past it (the sorted-copy path)
    box, pts = data.make_random_system(16, 3000, seed=8)
    pts = np.asarray(pts, np.float32)
    nq = freud_amd.AABBQuery(box, pts)
    nl = nq.query(pts, dict(r_max=r_max, exclude_ii=True)).toNeighborList()
    host = _host_nlist(nl)
    print(f"r_max {r_max}: longest segment {int(np.asarray(nl.neighbor_counts).max())}", flush=True)
    f = FilterSANN(True).compute(nq, nl)
    ref, unfilled = oracle.filter_sann(host, len(pts))
    _assert_filtered_equal(f.filtered_nlist, ref)
    for closed in (True, False):
        radii = np.linspace(0.8, 1.2, len(pts)).astype(np.float32)
        f = FilterRAD(True, closed).compute(nq, nl, points_radii=radii)
        ref, _ = oracle.filter_rad(_bt(box), pts, pts, host, terminate_after_blocked=closed, points_radii=radii)
        _assert_filtered_equal(f.filtered_nlist, ref)


def test_duplicate_bonds(oracle):
    # a list holding some bonds twice (same i, j, d; weights 1 and 2, or both
    # 1): equal (d, j) keys rank by weight, then by position, as the oracle's
    # (i, d, j, w) order does
    box, pts = data.make_random_system(12, 600, seed=12)
    pts = np.asarray(pts, np.float32)
    nq = freud_amd.AABBQuery(box, pts)
    nl = nq.query(pts, dict(r_max=1.8, exclude_ii=True)).toNeighborList()
    qi, pj, vec = np.asarray(nl.query_point_indices), np.asarray(nl.point_indices), np.asarray(nl.vectors)
    rng = np.random.default_rng(3)
    dup = rng.random(len(qi)) < 0.15
    rep = np.where(dup, 2, 1)
    w = np.ones(int(rep.sum()), np.float32)
    idx = np.repeat(np.arange(len(qi)), rep)
    second = np.concatenate([[False], idx[1:] == idx[:-1]])
    w[second] = np.where(rng.random(int(second.sum())) < 0.5, 2.0, 1.0)
    given = NeighborList.from_arrays(len(pts), len(pts), qi[idx], pj[idx], vec[idx], w)
    host = _host_nlist(given)
    print(f"duplicates: {int(dup.sum())} of {len(qi)} bonds doubled", flush=True)
    f = FilterSANN(True).compute(nq, given)
    ref, _ = oracle.filter_sann(host, len(pts))
    _assert_filtered_equal(f.filtered_nlist, ref)
    for closed in (True, False):
        f = FilterRAD(True, closed).compute(nq, given)
        ref, _ = oracle.filter_rad(_bt(box), pts, pts, host, terminate_after_blocked=closed)
        _assert_filtered_equal(f.filtered_nlist, ref)
