"""Potentials of mean force and torque: the API of ``freud.pmft`` (PMFTXY,
PMFTXYT, PMFTR12, PMFTXYZ; freud/pmft.py, freud/pmft/*.cc) on the GPU.

Every bond of the query (or of a given NeighborList) is mapped to its
coordinates in the query point's frame and binned into an integer histogram on
the device (``fa_pmft*_accumulate``); counts accumulate across ``compute``
calls until ``reset``.  The pair correlation function and the PMFT follow the
reference's normalisation (PMFT.h:76-87, PMFTXYZ.cc:92-103) in its float order.
"""

import ctypes

import numpy as np

from freud_amd import _lib
from freud_amd._lib import DeviceArray, check, lib
from freud_amd.locality import NeighborQuery, _SpatialHistogram
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


_TWO_PI = np.float32(2.0 * np.pi)


class _PMFT3(_PMFT):
    """Three-axis PMFTs: device counts, host reduction (``fa_pmft_reduce``)."""

    _kind = None

    def _init_axes(self, axes):
        # axes: [(n, min, max)] as RegularAxis(n, min, max)
        self._axes = [(int(n), np.float32(lo), np.float32(hi)) for n, lo, hi in axes]
        self._n = tuple(a[0] for a in self._axes)
        self._counts = None
        self._frames = 0

    def _reset(self):
        total = self._n[0] * self._n[1] * self._n[2]
        if self._counts is None:
            self._counts = DeviceArray.zeros(total, np.uint32)
        else:
            self._counts.zero()
        self._frames = 0
        self._reduced = None

    def _finish(self, nq, num_query_points):
        # BondHistogramCompute.h:115-126
        self._frames += 1
        self._n_points = nq._n
        self._n_query_points = int(num_query_points)
        self._box = nq._box
        self._reduced = None
        return self

    def _maxes(self):
        return np.array([0.0, 0.0, 0.0], np.float32)

    def _n_equiv(self):
        return 0

    def _reduce(self):
        if self._reduced is None:
            counts = self._counts.numpy()
            pcf = np.empty(counts.shape, np.float32)
            cbox = _lib.make_box(self._box)
            nb = np.array(self._n, np.uint32)
            mx = self._maxes()
            check(lib().fa_pmft_reduce(self._kind, ctypes.byref(cbox), counts.ctypes.data, nb.ctypes.data,
                                       mx.ctypes.data, self._n_points, self._n_query_points, self._frames,
                                       self._n_equiv(), pcf.ctypes.data))
            self._reduced = (counts.reshape(self._n), pcf.reshape(self._n))
        return self._reduced

    @computed_property
    def bin_counts(self):
        return self._reduce()[0].copy()

    @computed_property
    def _pcf(self):
        return self._reduce()[1].copy()

    @computed_property
    def box(self):
        import freud_amd.box

        return freud_amd.box.Box.from_box(self._box)

    @property
    def bin_edges(self):
        """RegularAxis edges min + float(i) * width per axis (Histogram.h:126-138)."""
        out = []
        for n, lo, hi in self._axes:
            width = np.float32((hi - lo) / np.float32(n))
            out.append((lo + np.arange(n + 1, dtype=np.float32) * width).astype(np.float32))
        return out

    @property
    def bin_centers(self):
        return [((e[:-1] + e[1:]) / np.float32(2.0)).astype(np.float32) for e in self.bin_edges]

    @property
    def bounds(self):
        return [(float(lo), float(hi)) for _, lo, hi in self._axes]

    @property
    def nbins(self):
        return list(self._n)


def _angles(orientations, n):
    return np.ascontiguousarray(_gen_angle_array(orientations, shape=(n,)), dtype=np.float32)


class PMFTXYT(_PMFT3):
    """PMFT in x, y and the relative orientation t (2D systems only;
    freud/pmft.py:224-323, PMFTXYT.cc)."""

    _kind = _lib.PMFT_XYT

    def __init__(self, x_max, y_max, bins):
        try:
            n_x, n_y, n_t = bins
        except TypeError:
            n_x = n_y = n_t = bins
        # PMFTXYT.cc:27-46
        for n, axis in ((n_x, "X"), (n_y, "Y"), (n_t, "T")):
            if n < 1:
                msg = f"PMFTXYT requires at least 1 bin in {axis}."
                raise ValueError(msg)
        if x_max < 0:
            msg = "PMFTXYT requires that x_max must be positive."
            raise ValueError(msg)
        if y_max < 0:
            msg = "PMFTXYT requires that y_max must be positive."
            raise ValueError(msg)
        self._init_axes([(n_x, -np.float32(x_max), x_max), (n_y, -np.float32(y_max), y_max), (n_t, 0, _TWO_PI)])
        self.r_max = np.sqrt(x_max**2 + y_max**2)

    def _maxes(self):
        return np.array([self._axes[0][2], self._axes[1][2], 0.0], np.float32)

    def compute(self, system, orientations, query_points=None, query_orientations=None, neighbors=None,
                reset=True):
        if reset or self._counts is None:
            self._reset()
        nq, nlist, qargs, d_q, num_query_points = self._preprocess_arguments(system, query_points, neighbors)
        if not nq._box.is2D:
            # PMFTXYT.cc:82 (Box::enforce2D)
            msg = "A 3D box was provided to a class that only supports 2D systems."
            raise ValueError(msg)
        ori = _angles(orientations, nq._n)
        qori = ori if query_orientations is None else _angles(query_orientations, num_query_points)
        nlist = self._neighbor_list(nq, nlist, qargs, d_q, num_query_points)
        check(lib().fa_pmftxyt_accumulate(ctypes.byref(nlist._c), ori.ctypes.data, qori.ctypes.data,
                                          float(self._axes[0][2]), float(self._axes[1][2]), self._n[0], self._n[1],
                                          self._n[2], self._counts.ptr))
        return self._finish(nq, num_query_points)

    def __repr__(self):
        return (f"freud_amd.pmft.PMFTXYT(x_max={float(self._axes[0][2])}, y_max={float(self._axes[1][2])}, "
                f"bins=({', '.join(str(b) for b in self._n)}))")


class PMFTR12(_PMFT3):
    """PMFT in r and the two bond angles t1, t2 (2D systems only;
    freud/pmft.py:124-221, PMFTR12.cc)."""

    _kind = _lib.PMFT_R12

    def __init__(self, r_max, bins):
        try:
            n_r, n_t1, n_t2 = bins
        except TypeError:
            n_r = n_t1 = n_t2 = bins
        # PMFTR12.cc:29-46
        for n, axis in ((n_r, "R"), (n_t1, "T1"), (n_t2, "T2")):
            if n < 1:
                msg = f"PMFTR12 requires at least 1 bin in {axis}."
                raise ValueError(msg)
        if r_max < 0:
            msg = "PMFTR12 requires that r_max must be positive."
            raise ValueError(msg)
        self._init_axes([(n_r, 0, r_max), (n_t1, 0, _TWO_PI), (n_t2, 0, _TWO_PI)])
        self.r_max = r_max

    def _maxes(self):
        return np.array([self._axes[0][2], 0.0, 0.0], np.float32)

    def compute(self, system, orientations, query_points=None, query_orientations=None, neighbors=None,
                reset=True):
        if reset or self._counts is None:
            self._reset()
        nq, nlist, qargs, d_q, num_query_points = self._preprocess_arguments(system, query_points, neighbors)
        if not nq._box.is2D:
            # PMFTR12.cc:96 (Box::enforce2D)
            msg = "A 3D box was provided to a class that only supports 2D systems."
            raise ValueError(msg)
        ori = _angles(orientations, nq._n)
        qori = ori if query_orientations is None else _angles(query_orientations, num_query_points)
        nlist = self._neighbor_list(nq, nlist, qargs, d_q, num_query_points)
        check(lib().fa_pmftr12_accumulate(ctypes.byref(nlist._c), ori.ctypes.data, qori.ctypes.data,
                                          float(self._axes[0][2]), self._n[0], self._n[1], self._n[2],
                                          self._counts.ptr))
        return self._finish(nq, num_query_points)

    def __repr__(self):
        return (f"freud_amd.pmft.PMFTR12(r_max={self.r_max}, bins=({', '.join(str(b) for b in self._n)}))")


class PMFTXYZ(_PMFT3):
    """PMFT in x, y, z in the query point's frame (3D systems only;
    freud/pmft.py:456-585, PMFTXYZ.cc)."""

    _kind = _lib.PMFT_XYZ

    def __init__(self, x_max, y_max, z_max, bins, shiftvec=None):
        if shiftvec is None:
            shiftvec = [0, 0, 0]
        try:
            n_x, n_y, n_z = bins
        except TypeError:
            n_x = n_y = n_z = bins
        # PMFTXYZ.cc:22-48
        for n, axis in ((n_x, "X"), (n_y, "Y"), (n_z, "Z")):
            if n < 1:
                msg = f"PMFTXYZ requires at least 1 bin in {axis}."
                raise ValueError(msg)
        for v, name in ((x_max, "x_max"), (y_max, "y_max"), (z_max, "z_max")):
            if v < 0:
                msg = f"PMFTXYZ requires that {name} must be positive."
                raise ValueError(msg)
        self._init_axes([(n_x, -np.float32(x_max), x_max), (n_y, -np.float32(y_max), y_max),
                         (n_z, -np.float32(z_max), z_max)])
        self.shiftvec = np.array(shiftvec, dtype=np.float32)
        self.r_max = np.sqrt(x_max**2 + y_max**2 + z_max**2)
        self._num_equiv = None

    def _reset(self):
        super()._reset()
        # PMFTXYZ.cc:99-103
        self._num_equiv = None

    def _maxes(self):
        return np.array([a[2] for a in self._axes], np.float32)

    def _n_equiv(self):
        return int(self._num_equiv)

    def compute(self, system, query_orientations, query_points=None, equiv_orientations=None, neighbors=None,
                reset=True):
        if reset or self._counts is None:
            self._reset()
        nq = NeighborQuery.from_system(system)
        if query_points is None:
            qp = nq.points
        else:
            qp = _convert_array(query_points, shape=(None, 3))
        # freud/pmft.py:549-551: the query points are shifted (a new array, so
        # the query takes them as distinct points; exclude_ii follows
        # query_points is None)
        shifted = np.ascontiguousarray(qp - self.shiftvec.reshape(1, 3), dtype=np.float32)
        nlist, qargs = self._resolve_neighbors(neighbors, query_points)
        num_query_points = shifted.shape[0]
        qori = _convert_array(np.atleast_1d(query_orientations), shape=(num_query_points, 4))
        if equiv_orientations is None:
            equiv = np.array([[1, 0, 0, 0]], dtype=np.float32)
        else:
            equiv = _convert_array(equiv_orientations, shape=(None, 4))
        n_equiv = equiv.shape[0]
        # PMFTXYZ.cc:114-124
        if self._num_equiv is None:
            self._num_equiv = n_equiv
        elif self._num_equiv != n_equiv:
            msg = "The number of equivalent orientations must be constant while accumulating data into PMFTXYZ."
            raise RuntimeError(msg)
        if nq._box.is2D:
            # PMFTXYZ.cc:124 (Box::enforce3D)
            msg = "A 2D box was provided to a class that only supports 3D systems."
            raise ValueError(msg)
        d_q = DeviceArray.from_numpy(shifted)
        if nlist is None:
            nlist = nq._run_query(d_q, num_query_points, qargs)
        nlist._validate(num_query_points, nq._n)
        check(lib().fa_pmftxyz_accumulate(ctypes.byref(nlist._c), qori.ctypes.data, equiv.ctypes.data, n_equiv,
                                          float(self._axes[0][2]), float(self._axes[1][2]),
                                          float(self._axes[2][2]), self._n[0], self._n[1], self._n[2],
                                          self._counts.ptr))
        return self._finish(nq, num_query_points)

    def __repr__(self):
        return (f"freud_amd.pmft.PMFTXYZ(x_max={float(self._axes[0][2])}, y_max={float(self._axes[1][2])}, "
                f"z_max={float(self._axes[2][2])}, bins=({', '.join(str(b) for b in self._n)}), "
                f"shiftvec={self.shiftvec.tolist()})")
