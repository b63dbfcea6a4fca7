"""Bond orientational order diagrams: the API of ``freud.environment.BondOrder``
(freud/environment.py:196-382, freud/environment/BondOrder.cc) on the GPU.

Every bond of a query (or of a given NeighborList) is turned into a direction
according to the mode (the bond itself, the bond in the point's frame, the
bond rotated by the relative orientation, or the neighbour's director) and
binned by its azimuthal and polar angles into an integer histogram on the
device (``fa_bond_order_accumulate``); counts accumulate across ``compute``
calls until ``reset``.  The bond order divides by each bin's surface area and
the number of frames (BondOrder.cc:68-74).
"""

import collections.abc
import ctypes

import numpy as np

from freud_amd._lib import DeviceArray, check, lib
from freud_amd.locality import _SpatialHistogram
from freud_amd.util import _convert_array, computed_property

_TWO_PI = np.float32(2.0 * np.pi)


class BondOrder(_SpatialHistogram):
    """Bond orientational order diagram (freud/environment.py:196-382)."""

    # BondOrder.h:23-29
    known_modes = {"bod": 0, "lbod": 1, "obcd": 2, "oocd": 3}

    def __init__(self, bins, mode="bod"):
        if isinstance(bins, collections.abc.Sequence):
            n_theta, n_phi = bins
        else:
            n_theta = n_phi = bins
        try:
            self._mode = self.known_modes[mode]
        except KeyError as err:
            msg = f"Unknown BondOrder mode: {mode}"
            raise ValueError(msg) from err
        # BondOrder.cc:16-24
        if n_theta < 2:
            msg = "BondOrder requires at least 2 bins in theta."
            raise ValueError(msg)
        if n_phi < 2:
            msg = "BondOrder requires at least 2 bins in phi."
            raise ValueError(msg)
        self._n = (int(n_theta), int(n_phi))
        self._mode_name = mode
        self._counts = None
        self._frames = 0

    @property
    def default_query_args(self):
        # freud/environment.py:286-292
        msg = ("The BondOrder class does not provide default query arguments. Query arguments "
               "or a neighbor list must be provided to this compute method.")
        raise NotImplementedError(msg)

    def _reset(self):
        if self._counts is None:
            self._counts = DeviceArray.zeros(self._n[0] * self._n[1], np.uint32)
        else:
            self._counts.zero()
        self._frames = 0
        self._reduced = None

    def compute(self, system, orientations=None, query_points=None, query_orientations=None, neighbors=None,
                reset=True):
        if reset or self._counts is None:
            self._reset()
        nq, nlist, qargs, d_q, num_query_points = self._preprocess_arguments(system, query_points, neighbors)
        # freud/environment.py:341-356
        if orientations is None:
            orientations = np.array([[1, 0, 0, 0]] * nq._n)
        if query_orientations is None:
            query_orientations = orientations
        orientations = _convert_array(orientations, shape=(nq._n, 4))
        query_orientations = _convert_array(query_orientations, shape=(num_query_points, 4))
        nlist = self._neighbor_list(nq, nlist, qargs, d_q, num_query_points)
        check(lib().fa_bond_order_accumulate(ctypes.byref(nlist._c), orientations.ctypes.data,
                                             query_orientations.ctypes.data, self._mode, self._n[0], self._n[1],
                                             self._counts.ptr))
        # BondHistogramCompute.h:115-126
        self._frames += 1
        self._box = nq._box
        self._reduced = None
        return self

    def _reduce(self):
        if self._reduced is None:
            counts = self._counts.numpy()
            bo = np.empty(counts.shape, np.float32)
            check(lib().fa_bond_order_reduce(counts.ctypes.data, self._n[0], self._n[1], self._frames,
                                             bo.ctypes.data))
            self._reduced = (counts.reshape(self._n), bo.reshape(self._n))
        return self._reduced

    @computed_property
    def bond_order(self):
        return self._reduce()[1].copy()

    @computed_property
    def bin_counts(self):
        return self._reduce()[0].copy()

    @computed_property
    def box(self):
        import freud_amd.box

        return freud_amd.box.Box.from_box(self._box)

    def _axes(self):
        return [(self._n[0], np.float32(0), _TWO_PI), (self._n[1], np.float32(0), np.float32(np.pi))]

    @property
    def bin_edges(self):
        out = []
        for n, lo, hi in self._axes():
            width = np.float32((hi - lo) / np.float32(n))
            out.append((lo + np.arange(n + 1, dtype=np.float32) * width).astype(np.float32))
        return out

    @property
    def bin_centers(self):
        return [((e[:-1] + e[1:]) / np.float32(2.0)).astype(np.float32) for e in self.bin_edges]

    @property
    def bounds(self):
        return [(float(lo), float(hi)) for _, lo, hi in self._axes()]

    @property
    def nbins(self):
        return list(self._n)

    @property
    def mode(self):
        return self._mode_name

    def __repr__(self):
        return f"freud_amd.environment.BondOrder(bins=({', '.join(str(b) for b in self._n)}), mode='{self.mode}')"
