"""Parity at the benchmark sizes, against the oracle on the same inputs.

* The large single frame: 1,048,576 jittered FCC points (64^3 cells x 4,
  a = 1.5874, sigma = 0.05 a, seed 1).  Ball query r_max = 1.5 and kNN k = 12
  NeighborLists bit-exact (AABBQuery.cc:77-281, NeighborQuery.h:434-481); RDF
  200 bins at r_max = 5 bin counts bit-exact (RDF.cc:101-110,
  Histogram.h:152-174) with g(r) and N(r) within 1e-6 relative; q6, w6 (plain
  and normalised) and averaged q6 within 1e-5 of the oracle's float32 run and
  of its float64 re-evaluation (Steinhardt.cc:120-359).
* The plain 4,096-point parity case on a LinkCell with cell_width 4 (4 cells
  per side): ball r_max = 4 NeighborList and RDF 100 bins bit-exact
  (LinkCell.cc:222-367, 496-573).
"""

import numpy as np
import numpy.testing as npt
import pytest

import freud_amd

pytestmark = pytest.mark.gpu

THREADS = 16


def _bt(box):
    return (box.Lx, box.Ly, box.Lz, box.xy, box.xz, box.yz, box.is2D)


def _assert_nlist_equal(nl, ref):
    npt.assert_array_equal(nl[:], ref["ij"])
    npt.assert_array_equal(nl.distances, ref["distances"])
    npt.assert_array_equal(nl.vectors, ref["vectors"])
    npt.assert_array_equal(nl.weights, 1.0)


@pytest.fixture(scope="module")
def frame():
    return freud_amd.data.fcc_jittered(64, 1.5874, 0.05 * 1.5874, seed=1)


@pytest.fixture(scope="module")
def ball(oracle, frame):
    box, pts = frame
    aq = freud_amd.AABBQuery(box, pts)
    nl = aq.query(pts, dict(r_max=1.5, exclude_ii=True)).toNeighborList()
    ref = oracle.query(_bt(box), pts, pts, dict(r_max=1.5, exclude_ii=True), nthreads=THREADS)
    return aq, nl, ref


def test_frame_1M_ball_query(ball, frame):
    _, nl, ref = ball
    assert len(frame[1]) == 1_048_576
    assert len(ref["ij"]) > 13_000_000
    _assert_nlist_equal(nl, ref)
    npt.assert_array_equal(nl.segments, np.searchsorted(ref["ij"][:, 0], np.arange(len(frame[1]))))


def test_frame_1M_knn(oracle, ball, frame):
    box, pts = frame
    aq = ball[0]
    nl = aq.query(pts, dict(num_neighbors=12, exclude_ii=True)).toNeighborList()
    ref = oracle.query(_bt(box), pts, pts, dict(num_neighbors=12, exclude_ii=True), nthreads=THREADS)
    assert len(ref["ij"]) == 12 * len(pts)
    _assert_nlist_equal(nl, ref)


def test_frame_1M_rdf(oracle, ball, frame):
    box, pts = frame
    aq = ball[0]
    rdf = freud_amd.density.RDF(200, 5.0).compute(aq)
    counts = oracle.rdf_accumulate(_bt(box), pts, pts, 200, 5.0, exclude_ii=True, nthreads=THREADS)
    assert counts.astype(np.int64).sum() > 5e8
    npt.assert_array_equal(rdf.bin_counts, counts)
    g, n_r = oracle.rdf_reduce(_bt(box), counts, 200, 0.0, 5.0, len(pts), len(pts), 1)
    npt.assert_allclose(rdf.rdf, g, rtol=1e-6, atol=0)
    npt.assert_allclose(rdf.n_r, n_r, rtol=1e-6, atol=0)


def test_frame_1M_steinhardt(oracle, ball, frame):
    box, pts = frame
    _, nl, ref = ball
    bt = _bt(box)
    errs = {}
    for kw, key in ((dict(), "ql"), (dict(wl=True), "wl"), (dict(wl=True, wl_normalize=True), "wl"),
                    (dict(average=True), "ql_ave")):
        st = freud_amd.order.Steinhardt(6, **kw).compute((box, pts), neighbors=nl)
        r32 = oracle.steinhardt(bt, pts, ref, [6], nthreads=THREADS, **kw)
        r64 = oracle.steinhardt(bt, pts, ref, [6], f64=True, nthreads=THREADS, **kw)
        got = np.asarray(st.particle_order, np.float64)
        e32 = float(np.max(np.abs(got - r32[key][:, 0])))
        e64 = float(np.max(np.abs(got - r64[key][:, 0])))
        errs[str(kw)] = (e32, e64)
        assert e32 <= 1e-5 and e64 <= 1e-5, (kw, e32, e64)
        npt.assert_allclose(st.order, r64["order"][0], atol=1e-5)
    print("max |error| against the oracle float32 / float64 runs:", errs)


def test_linkcell_width4_parity(oracle):
    rng = np.random.default_rng(0)
    n, L = 4096, 16.0
    pts = ((rng.random((n, 3)) - 0.5) * L).astype(np.float32)
    box = freud_amd.Box.cube(L)
    lc = freud_amd.LinkCell(box, pts, cell_width=4.0)
    assert lc.cell_width == 4.0
    nl = lc.query(pts, dict(r_max=4.0, exclude_ii=True)).toNeighborList()
    ref = oracle.query(_bt(box), pts, pts, dict(r_max=4.0, exclude_ii=True), nq_type="linkcell", cell_width=4.0,
                       nthreads=THREADS)
    assert len(ref["ij"]) > 1_000_000
    _assert_nlist_equal(nl, ref)
    rdf = freud_amd.density.RDF(100, 4.0).compute(lc)
    counts = oracle.rdf_accumulate(_bt(box), pts, pts, 100, 4.0, exclude_ii=True, nq_type="linkcell",
                                   cell_width=4.0, nthreads=THREADS)
    npt.assert_array_equal(rdf.bin_counts, counts)
    # the default cell width as well
    lc0 = freud_amd.LinkCell(box, pts)
    nl0 = lc0.query(pts, dict(r_max=4.0, exclude_ii=True)).toNeighborList()
    _assert_nlist_equal(nl0, oracle.query(_bt(box), pts, pts, dict(r_max=4.0, exclude_ii=True), nq_type="linkcell",
                                          nthreads=THREADS))
