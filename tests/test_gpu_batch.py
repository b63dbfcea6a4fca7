"""GPU parity of the batched trajectory path: every frame of a batch against
the oracle on the same inputs (NeighborList and RDF counts bit for bit, q_l
within 1e-5), eager and graph-replayed runs, new positions after capture, and
a replay whose bonds outgrow the allocation."""

import numpy as np
import numpy.testing as npt
import pytest

import freud_amd
from freud_amd.box import Box
from freud_amd.batch import TrajectoryBatch

pytestmark = pytest.mark.gpu

R_BALL, BINS, R_RDF, L_Q = 1.5, 64, 3.6, 6


def _bt(box):
    return (box.Lx, box.Ly, box.Lz, box.xy, box.xz, box.yz, box.is2D)


def _frames(seed=3):
    """Mixed frames: liquid and crystals, cubic/triclinic, a 2D frame, a frame
    with points outside its box, and sizes from 1 to a few thousand."""
    rng = np.random.default_rng(seed)
    out = []
    fr, _ = freud_amd.data.mixed_frames(4, 700, 3000, seed=seed)
    out += fr
    box = Box(9.0, 10.0, 9.5, 0.2, -0.25, 0.1)
    out.append((box, box.make_absolute(rng.random((900, 3)).astype(np.float32))))
    b2 = Box.square(14.0)
    p2 = b2.make_absolute(rng.random((600, 3)).astype(np.float32))
    p2[:, 2] = 0
    out.append((b2, p2))
    b3 = Box.cube(10.0)
    p3 = ((rng.random((800, 3)) - 0.5) * 10).astype(np.float32)
    p3[:40] *= 1.04  # a few points just outside the box
    out.append((b3, p3))
    out.append((Box.cube(8.0), np.zeros((1, 3), np.float32)))
    return out


def _check(batch, frames, oracle):
    counts = batch.rdf_counts
    qls = batch.ql()
    order = batch.order
    nb = batch.num_bonds
    for f, (box, pts) in enumerate(frames):
        ref = oracle.query(_bt(box), pts, pts, dict(r_max=R_BALL, exclude_ii=True), nthreads=8)
        nl = batch.neighbor_list(f)
        npt.assert_array_equal(nl[:], ref["ij"], err_msg=f"frame {f}")
        npt.assert_array_equal(nl.distances, ref["distances"], err_msg=f"frame {f}")
        npt.assert_array_equal(nl.vectors, ref["vectors"], err_msg=f"frame {f}")
        assert int(nb[f]) == len(ref["ij"])
        c = np.bincount(ref["ij"][:, 0], minlength=len(pts)) if len(ref["ij"]) else np.zeros(len(pts))
        npt.assert_array_equal(nl.neighbor_counts, c)
        rc = oracle.rdf_accumulate(_bt(box), pts, pts, BINS, R_RDF, exclude_ii=True, nthreads=8)
        npt.assert_array_equal(counts[f], rc, err_msg=f"frame {f}")
        st = oracle.steinhardt(_bt(box), pts, ref, [L_Q], nthreads=8)
        ok = np.isfinite(st["ql"][:, 0])
        npt.assert_array_equal(np.isfinite(qls[f]), ok)
        npt.assert_allclose(qls[f][ok], st["ql"][ok, 0], atol=1e-5, err_msg=f"frame {f}")
        if ok.all():
            npt.assert_allclose(order[f], st["order"][0], rtol=1e-5, atol=1e-6)


def _make(frames):
    return TrajectoryBatch(frames, ball_r_max=R_BALL, rdf_bins=BINS, rdf_r_max=R_RDF, l=L_Q)


def test_batch_eager_matches_oracle(oracle):
    frames = _frames()
    b = _make(frames).compute(use_graph=False)
    _check(b, frames, oracle)


def test_batch_graph_replay_and_new_points(oracle):
    frames = _frames(seed=5)
    b = _make(frames)
    b.compute(use_graph=True)  # eager run, then capture
    assert b.config["graph"]
    first = b.rdf_counts.copy()
    b.compute(use_graph=True)  # replay
    npt.assert_array_equal(b.rdf_counts, first)
    _check(b, frames, oracle)
    # new positions through the same graph
    rng = np.random.default_rng(9)
    moved = []
    for box, pts in frames:
        p = pts + rng.normal(0, 0.05, pts.shape).astype(np.float32)
        if not box.is2D:
            p = box.wrap(p)
        else:
            p[:, 2] = 0
            p = box.wrap(p)
        moved.append((box, np.ascontiguousarray(p, np.float32)))
    b.set_points([p for _, p in moved]).compute(use_graph=True)
    _check(b, moved, oracle)


def test_batch_bond_overflow_reruns(oracle):
    rng = np.random.default_rng(2)
    box = Box.cube(12.0)
    pts = box.make_absolute(rng.random((1500, 3)).astype(np.float32))
    frames = [(box, pts), (Box.cube(11.0), Box.cube(11.0).make_absolute(rng.random((1200, 3)).astype(np.float32)))]
    b = _make(frames)
    b.compute(use_graph=True)
    cap = b.config["bond_capacity"]
    # squeeze frame 0 into a corner: far more bonds than the allocation holds
    dense = (rng.random((1500, 3)) * 3.0 - 6.0).astype(np.float32)
    squeezed = [(box, dense), frames[1]]
    b.set_points([p for _, p in squeezed]).compute(use_graph=True)
    assert b.config["bond_capacity"] > cap
    _check(b, squeezed, oracle)
    # and back through the re-captured graph
    b.set_points([p for _, p in frames]).compute(use_graph=True)
    _check(b, frames, oracle)


def test_batch_dense_ball_query(oracle):
    """A ball radius at which the frames expect more bonds per point than the
    walk's key stage holds (r_max 2.2 at density ~1: ~45 bonds): the batch
    counts and fills instead of staging keys.  NeighborLists against the
    oracle, eager and graph-replayed."""
    r = 2.2
    frames = _frames(seed=5)
    b = TrajectoryBatch(frames, ball_r_max=r, rdf_bins=BINS, rdf_r_max=R_RDF, l=L_Q)
    for use_graph in (False, True):
        b.compute(use_graph=use_graph)
        nb = b.num_bonds
        for f, (box, pts) in enumerate(frames):
            ref = oracle.query(_bt(box), pts, pts, dict(r_max=r, exclude_ii=True), nthreads=8)
            nl = b.neighbor_list(f)
            npt.assert_array_equal(nl[:], ref["ij"], err_msg=f"frame {f}")
            npt.assert_array_equal(nl.distances, ref["distances"], err_msg=f"frame {f}")
            npt.assert_array_equal(nl.vectors, ref["vectors"], err_msg=f"frame {f}")
            assert int(nb[f]) == len(ref["ij"])


def test_batch_matches_single_frame_api():
    frames = _frames(seed=7)
    b = _make(frames).compute()
    for f, (box, pts) in enumerate(frames):
        aq = freud_amd.AABBQuery(box, pts)
        nl = aq.query(pts, dict(r_max=R_BALL, exclude_ii=True)).toNeighborList()
        bn = b.neighbor_list(f)
        npt.assert_array_equal(bn[:], nl[:])
        npt.assert_array_equal(bn.segments, nl.segments)
        rdf = freud_amd.density.RDF(BINS, R_RDF).compute(aq)
        npt.assert_array_equal(b.rdf_counts[f], rdf.bin_counts)
        st = freud_amd.order.Steinhardt(L_Q).compute(aq, neighbors=nl)
        npt.assert_allclose(b.ql()[f], st.particle_order, atol=1e-6, equal_nan=True)


def test_batch_rdf_every_launch_shape(oracle):
    """Each batched RDF launch shape the autotuner can pick ((R, Fx) = (2, 1),
    (3, 1), (4, 1), (2, 4), (2, 8): cells r/R wide, Fx times narrower along x)
    gives the oracle's counts in every frame."""
    from freud_amd import _lib

    frames = _frames(seed=11)
    pts_total = sum(len(p) for _, p in frames)
    vol = sum(b.volume for b, _ in frames)
    key = int(np.floor(2.0 * np.log2(pts_total / vol * R_RDF ** 3) + 0.5)) + 1000
    ref = [oracle.rdf_accumulate(_bt(box), pts, pts, BINS, R_RDF, exclude_ii=True, nthreads=8)
           for box, pts in frames]
    before = _lib.tune_table()
    try:
        for shape in range(5):
            _lib.tune_load({(1, key): shape})
            b = _make(frames).compute(use_graph=False)
            cfg = b.config
            assert (cfg["rdf_R"], cfg["rdf_F"]) == ((2, 1), (3, 1), (4, 1), (2, 4), (2, 8))[shape]
            counts = b.rdf_counts
            for f in range(len(frames)):
                npt.assert_array_equal(counts[f], ref[f], err_msg=f"shape {shape} frame {f}")
    finally:
        _lib.tune_load({(1, key): before.get((1, key), -1)})


def test_batched_trajectory_full_size(oracle):
    """The batched trajectory at its benchmark size: 256 frames of 10k-200k
    points (log-uniform; liquid, jittered FCC and BCC; cubic and triclinic),
    ball r_max = 1.5, RDF 200 bins to r_max = 5, q6, run eagerly and then
    replayed from the captured graph.  Every frame's RDF counts and bond count
    equal the oracle's; the NeighborLists and q6 of the smallest, the largest
    and a triclinic frame are checked in full."""
    frames, kinds = freud_amd.data.mixed_frames(256, 10_000, 200_000, seed=4)
    sizes = np.array([len(p) for _, p in frames])
    assert sizes.min() >= 9_000 and sizes.max() <= 220_000 and sizes.sum() > 1.2e7
    b = TrajectoryBatch(frames, ball_r_max=1.5, rdf_bins=200, rdf_r_max=5.0, l=6)
    b.compute(use_graph=True)  # eager run, then capture
    eager = b.rdf_counts.copy()
    b.compute(use_graph=True)  # graph replay
    assert b.config["graph"]
    counts = b.rdf_counts
    npt.assert_array_equal(counts, eager)
    nb = b.num_bonds
    for f, (box, pts) in enumerate(frames):
        rc = oracle.rdf_accumulate(_bt(box), pts, pts, 200, 5.0, exclude_ii=True, nthreads=16)
        npt.assert_array_equal(counts[f], rc, err_msg=f"frame {f}")
    tri = [f for f, (box, _) in enumerate(frames) if box.xy != 0 or box.xz != 0 or box.yz != 0]
    assert tri
    sample = {int(np.argmin(sizes)), int(np.argmax(sizes)), tri[0]}
    qls = b.ql()
    for f, (box, pts) in enumerate(frames):
        ref = oracle.query(_bt(box), pts, pts, dict(r_max=1.5, exclude_ii=True), nthreads=16)
        assert int(nb[f]) == len(ref["ij"]), f"frame {f}"
        if f not in sample:
            continue
        nl = b.neighbor_list(f)
        npt.assert_array_equal(nl[:], ref["ij"], err_msg=f"frame {f}")
        npt.assert_array_equal(nl.distances, ref["distances"], err_msg=f"frame {f}")
        npt.assert_array_equal(nl.vectors, ref["vectors"], err_msg=f"frame {f}")
        st = oracle.steinhardt(_bt(box), pts, ref, [6], nthreads=16)
        ok = np.isfinite(st["ql"][:, 0])
        npt.assert_array_equal(np.isfinite(qls[f]), ok)
        npt.assert_allclose(qls[f][ok], st["ql"][ok, 0], atol=1e-5, err_msg=f"frame {f}")
