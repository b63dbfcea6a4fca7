"""Potential of mean force and torque in 2D Cartesian coordinates: the API of
``freud.pmft.PMFTXY`` (freud/pmft.py:328-453, freud/pmft/PMFTXY.cc) on the GPU.

Every bond vector of the query (or of a given NeighborList) is rotated into
its query point's frame and binned into an integer (n_x, n_y) histogram on the
device (``fa_pmftxy_accumulate``); counts accumulate across ``compute`` calls
until ``reset``.  The pair correlation function and the PMFT follow the
reference's normalisation (PMFT.h:76-87) in its float order.
"""

import ctypes

import numpy as np

from freud_amd import _lib
from freud_amd._lib import DeviceArray, check, lib
from freud_amd.locality import _SpatialHistogram
from freud_amd.util import _convert_array, computed_property


def _quat_to_z_angle(orientations, num_points):
    """Quaternions (w, x, y, z) to rotation angles about z (freud/pmft.py:57-81):
    an array is taken as quaternions when its last axis has length 4 (a
    1D array of length 4 only when there are not exactly 4 points).  Their
    rotation angle is 2 atan2(|v|, w); the rotations must be about +z (or
    all the identity)."""
    o = np.asarray(orientations)
    is_quat = (o.ndim == 1 and o.shape[0] == 4 and num_points != 4) or (o.ndim == 2 and o.shape[1] == 4)
    if not is_quat:
        return o
    q = np.atleast_2d(o).astype(np.float64)
    norm_v = np.linalg.norm(q[:, 1:], axis=-1)
    angles = 2.0 * np.arctan2(norm_v, q[:, 0])
    with np.errstate(invalid="ignore", divide="ignore"):
        axes = np.where(norm_v[:, None] > 0, q[:, 1:] / norm_v[:, None], np.array([1.0, 0.0, 0.0]))
    if not (np.allclose(angles, 0) or np.allclose(axes, [0, 0, 1])):
        msg = "Orientations provided as quaternions must represent rotations about the z-axis."
        raise ValueError(msg)
    return angles if o.ndim == 2 else angles[0]


def _gen_angle_array(orientations, shape):
    """freud/pmft.py:84-94."""
    return _convert_array(np.atleast_1d(_quat_to_z_angle(np.asarray(orientations).squeeze(), shape[0])),
                          shape=shape)


class _PMFT(_SpatialHistogram):
    @computed_property
    def pmft(self):
        """-ln(pcf); +inf where no bond fell into a bin (freud/pmft.py:111-116)."""
        with np.errstate(divide="ignore"):
            return -np.log(np.copy(self._pcf))


class PMFTXY(_PMFT):
    """PMFT in x, y (2D systems only)."""

    def __init__(self, x_max, y_max, bins):
        try:
            n_x, n_y = bins
        except TypeError:
            n_x = n_y = bins
        # PMFTXY.cc:25-41
        if n_x < 1:
            msg = "PMFTXY requires at least 1 bin in X."
            raise ValueError(msg)
        if n_y < 1:
            msg = "PMFTXY requires at least 1 bin in Y."
            raise ValueError(msg)
        if x_max < 0:
            msg = "PMFTXY requires that x_max must be positive."
            raise ValueError(msg)
        if y_max < 0:
            msg = "PMFTXY requires that y_max must be positive."
            raise ValueError(msg)
        self._n = (int(n_x), int(n_y))
        self._max32 = (np.float32(x_max), np.float32(y_max))
        self.r_max = np.sqrt(x_max**2 + y_max**2)
        self._counts = None
        self._frames = 0

    def _reset(self):
        if self._counts is None:
            self._counts = DeviceArray.zeros(self._n[0] * self._n[1], np.uint32)
        else:
            self._counts.zero()
        self._frames = 0
        self._reduced = None

    def compute(self, system, query_orientations, query_points=None, neighbors=None, reset=True):
        if reset or self._counts is None:
            self._reset()
        nq, nlist, qargs, d_q, num_query_points = self._preprocess_arguments(system, query_points, neighbors)
        if not nq._box.is2D:
            # PMFTXY.cc:74 (Box::enforce2D)
            msg = "A 3D box was provided to a class that only supports 2D systems."
            raise ValueError(msg)
        angles = np.ascontiguousarray(_gen_angle_array(query_orientations, shape=(num_query_points,)),
                                      dtype=np.float32)
        nlist = self._neighbor_list(nq, nlist, qargs, d_q, num_query_points)
        check(lib().fa_pmftxy_accumulate(ctypes.byref(nlist._c), angles.ctypes.data, float(self._max32[0]),
                                         float(self._max32[1]), self._n[0], self._n[1], self._counts.ptr))
        # BondHistogramCompute.h:115-126
        self._frames += 1
        self._n_points = nq._n
        self._n_query_points = int(num_query_points)
        self._box = nq._box
        self._reduced = None
        return self

    def _reduce(self):
        if self._reduced is None:
            counts = self._counts.numpy()
            pcf = np.empty(counts.shape, np.float32)
            cbox = _lib.make_box(self._box)
            check(lib().fa_pmftxy_reduce(ctypes.byref(cbox), counts.ctypes.data, self._n[0], self._n[1],
                                         float(self._max32[0]), float(self._max32[1]), self._n_points,
                                         self._n_query_points, self._frames, pcf.ctypes.data))
            self._reduced = (counts.reshape(self._n), pcf.reshape(self._n))
        return self._reduced

    @computed_property
    def bin_counts(self):
        return np.squeeze(self._reduce()[0].copy())

    @computed_property
    def _pcf(self):
        return self._reduce()[1].copy()

    @computed_property
    def box(self):
        import freud_amd.box

        return freud_amd.box.Box.from_box(self._box)

    def _axis_edges(self, k):
        """RegularAxis(n, -max, max) edges: min + float(i) * width (Histogram.h:126-138)."""
        vmax = self._max32[k]
        vmin = np.float32(-vmax)
        width = np.float32((vmax - vmin) / np.float32(self._n[k]))
        return (vmin + np.arange(self._n[k] + 1, dtype=np.float32) * width).astype(np.float32)

    @property
    def bin_edges(self):
        return [self._axis_edges(0), self._axis_edges(1)]

    @property
    def bin_centers(self):
        out = []
        for e in self.bin_edges:
            out.append(((e[:-1] + e[1:]) / np.float32(2.0)).astype(np.float32))
        return out

    @property
    def bounds(self):
        return [(-float(self._max32[0]), float(self._max32[0])), (-float(self._max32[1]), float(self._max32[1]))]

    @property
    def nbins(self):
        return list(self._n)

    def __repr__(self):
        return (f"freud_amd.pmft.PMFTXY(x_max={float(self._max32[0])}, y_max={float(self._max32[1])}, "
                f"bins=({self._n[0]}, {self._n[1]}))")
