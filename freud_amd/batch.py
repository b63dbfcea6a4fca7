"""Batched trajectories: many frames analysed in one pass on the GPU.

The reference analyses a trajectory one frame at a time, e.g.::

    rdf = freud.density.RDF(bins=200, r_max=5)
    for box, points in frames:
        rdf.compute((box, points), reset=False)

(freud/density.py:635-647 per frame).  :class:`TrajectoryBatch` runs the ball
query, the RDF and Steinhardt q_l of every frame of a batch together (about
twenty kernel launches for the whole batch instead of that many per frame,
replayed from a HIP graph after the first run).  Per frame, the results are the
single-frame ones: the ball-query NeighborList and RDF bin counts bit for bit,
q_l per particle and the frame's system order.
"""

import ctypes

import numpy as np

from freud_amd import _lib
from freud_amd._lib import check, lib
from freud_amd.box import Box
from freud_amd.locality import NeighborList


class TrajectoryBatch:
    """A batch of frames of any box shape and size.

    Args:
        frames: sequence of ``(box, points)`` with float32 ``(N_f, 3)`` points
            (boxes may differ: cubic, triclinic or 2D).
        ball_r_max (float): ball-query radius (0 to skip the query and q_l).
        rdf_bins (int): RDF bins over ``[0, rdf_r_max)`` (0 to skip).
        rdf_r_max (float): RDF cutoff.
        l (int or None): Steinhardt l over the ball bonds (None to skip).
    """

    def __init__(self, frames, ball_r_max=1.5, rdf_bins=200, rdf_r_max=5.0, l=6):
        boxes, sizes = [], []
        for box, pts in frames:
            box = box if isinstance(box, Box) else Box.from_box(box)
            boxes.append(box)
            sizes.append(int(np.asarray(pts).shape[0]))
        self._boxes = boxes
        self._sizes = np.asarray(sizes, np.int64)
        self._offsets = np.concatenate([[0], np.cumsum(self._sizes)]).astype(np.uint64)
        self._n_frames = len(boxes)
        self._bins = int(rdf_bins)
        self._l = None if l is None else int(l)
        self._ball = float(ball_r_max)
        cboxes = (_lib.fa_box * max(1, self._n_frames))()
        for f, box in enumerate(boxes):
            cboxes[f] = _lib.make_box(box)
        plan = _lib.fa_batch_plan(float(ball_r_max), int(rdf_bins), float(rdf_r_max),
                                  0xFFFFFFFF if l is None else int(l))
        h = ctypes.c_void_p()
        check(lib().fa_batch_create(self._n_frames, ctypes.cast(cboxes, ctypes.c_void_p),
                                    self._offsets.ctypes.data, ctypes.byref(plan), ctypes.byref(h)))
        self._h = h
        self.set_points([p for _, p in frames])

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib._lib is not None:
            _lib._lib.fa_batch_destroy(h)

    @property
    def n_frames(self):
        return self._n_frames

    @property
    def sizes(self):
        return self._sizes.copy()

    def points_ptr(self):
        """Device address of the batch's (total, 3) float32 point array."""
        p = ctypes.c_void_p()
        check(lib().fa_batch_points(self._h, ctypes.byref(p)))
        return p.value

    def set_points(self, points):
        """Load new positions (same frame sizes) into the batch."""
        if len(points) != self._n_frames:
            raise ValueError("one point array per frame is required")
        allp = np.empty((int(self._offsets[-1]), 3), np.float32)
        for f, p in enumerate(points):
            p = np.asarray(p, np.float32).reshape(-1, 3)
            if p.shape[0] != self._sizes[f]:
                raise ValueError("frame sizes are fixed when the batch is created")
            allp[int(self._offsets[f]):int(self._offsets[f + 1])] = p
        if allp.size:
            check(lib().fa_memcpy_h2d(self.points_ptr(), allp.ctypes.data, allp.nbytes))
        return self

    def compute(self, use_graph=True):
        """Run every planned analysis over all frames (synchronises)."""
        check(lib().fa_batch_run(self._h, int(bool(use_graph))))
        return self

    @property
    def num_bonds(self):
        out = np.zeros(self._n_frames, np.uint64)
        check(lib().fa_batch_num_bonds(self._h, out.ctypes.data))
        return out

    def neighbor_list(self, frame):
        """Frame ``frame``'s ball-query NeighborList (frame-local indices)."""
        nl = NeighborList()
        check(lib().fa_batch_nlist(self._h, int(frame), ctypes.byref(nl._c)))
        return nl

    @property
    def rdf_counts(self):
        """(n_frames, bins) uint32 bin counts of the last run."""
        p = ctypes.c_void_p()
        check(lib().fa_batch_rdf_counts(self._h, ctypes.byref(p)))
        return _lib.copy_d2h(p.value, (self._n_frames, self._bins), np.uint32)

    def ql(self):
        """List of per-frame q_l arrays of the last run."""
        p = ctypes.c_void_p()
        check(lib().fa_batch_ql(self._h, ctypes.byref(p), None, None))
        allq = _lib.copy_d2h(p.value, (int(self._offsets[-1]),), np.float32)
        return [allq[int(self._offsets[f]):int(self._offsets[f + 1])] for f in range(self._n_frames)]

    def qlm(self):
        """List of per-frame (N_f, 2l+1) complex64 q_lm arrays of the last run."""
        p = ctypes.c_void_p()
        check(lib().fa_batch_ql(self._h, None, ctypes.byref(p), None))
        nm = 2 * self._l + 1
        a = _lib.copy_d2h(p.value, (int(self._offsets[-1]), nm, 2), np.float32)
        a = a[..., 0] + 1j * a[..., 1]
        return [a[int(self._offsets[f]):int(self._offsets[f + 1])].astype(np.complex64)
                for f in range(self._n_frames)]

    @property
    def order(self):
        """(n_frames,) system q_l of each frame."""
        out = np.zeros(max(1, self._n_frames), np.float32)
        check(lib().fa_batch_ql(self._h, None, None, out.ctypes.data))
        return out[:self._n_frames]

    @property
    def config(self):
        """Launch shape chosen by the autotuner: ball and RDF cells r/R wide (RDF
        cells rdf_F times narrower along x)."""
        a, b, f, g = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        cap = ctypes.c_uint64()
        check(lib().fa_batch_config(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(f), ctypes.byref(cap),
                                    ctypes.byref(g)))
        return {"ball_R": a.value, "rdf_R": b.value, "rdf_F": f.value, "bond_capacity": int(cap.value),
                "graph": bool(g.value)}
