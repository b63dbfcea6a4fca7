"""Radial distribution function: the API of ``freud.density.RDF``
(freud/density.py:535-696) on the GPU.

Without a NeighborList the bins are filled by the fused query + histogram
kernel (``fa_rdf_accumulate_query``); a given NeighborList is binned from its
distances.  Counts stay on the device across frames (``reset=False``), and the
normalisation (RDF.cc:73-99) runs in the reference's float order.
"""

import ctypes

import numpy as np

from freud_amd import _lib
from freud_amd._lib import DeviceArray, check, lib
from freud_amd.locality import _SpatialHistogram1D
from freud_amd.util import _Compute, computed_property


class RDF(_SpatialHistogram1D):
    """g(r) and N(r) (freud/density/RDF.cc)."""

    def __init__(self, bins, r_max, r_min=0, normalization_mode="exact"):
        # RDF.cc:27-42
        if bins == 0:
            msg = "RDF requires a nonzero number of bins."
            raise ValueError(msg)
        if r_max <= 0:
            msg = "RDF requires r_max to be positive."
            raise ValueError(msg)
        if r_min < 0:
            msg = "RDF requires r_min to be non-negative."
            raise ValueError(msg)
        if r_max <= r_min:
            msg = "RDF requires that r_max must be greater than r_min."
            raise ValueError(msg)
        self._bins = int(bins)
        self._r_min32 = np.float32(r_min)
        self._r_max32 = np.float32(r_max)
        self.mode = normalization_mode
        self.r_max = r_max
        self._counts = None
        self._frames = 0
        self._n_points = 0
        self._n_query_points = 0
        self._box = None
        self._called_compute = False

    @property
    def mode(self):
        return self._mode

    @mode.setter
    def mode(self, value):
        if value not in ("exact", "finite_size"):
            msg = "An invalid normalization mode was provided."
            raise ValueError(msg)
        self._mode = value

    def _reset(self):
        if self._counts is None:
            self._counts = DeviceArray.zeros(self._bins, np.uint32)
        else:
            self._counts.zero()
        self._frames = 0

    def compute(self, system, query_points=None, neighbors=None, reset=True):
        if reset or self._counts is None:
            self._reset()
        nq, nlist, qargs, d_q, num_query_points = self._preprocess_arguments(system, query_points, neighbors)
        if nlist is not None:
            check(lib().fa_rdf_accumulate_nlist(ctypes.byref(nlist._c), self._bins, float(self._r_min32),
                                                float(self._r_max32), self._counts.ptr))
        elif qargs.mode in ("ball", None) and qargs.num_neighbors == 0xFFFFFFFF:
            r_max = qargs.r_max
            check(lib().fa_rdf_accumulate_query(nq._handle, d_q.ptr, int(num_query_points), float(r_max),
                                                float(qargs.r_min), int(qargs.exclude_ii), self._bins,
                                                float(self._r_min32), float(self._r_max32), self._counts.ptr))
        else:
            nl = nq._run_query(d_q, num_query_points, qargs)
            check(lib().fa_rdf_accumulate_nlist(ctypes.byref(nl._c), self._bins, float(self._r_min32),
                                                float(self._r_max32), self._counts.ptr))
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
            rdf = np.empty(self._bins, np.float32)
            n_r = np.empty(self._bins, np.float32)
            cbox = _lib.make_box(self._box)
            check(lib().fa_rdf_reduce(ctypes.byref(cbox), counts.ctypes.data, self._bins, float(self._r_min32),
                                      float(self._r_max32), self._n_points, self._n_query_points, self._frames,
                                      int(self._mode == "finite_size"), rdf.ctypes.data, n_r.ctypes.data))
            self._reduced = (counts, rdf, n_r)
        return self._reduced

    @computed_property
    def rdf(self):
        return self._reduce()[1].copy()

    @computed_property
    def n_r(self):
        return self._reduce()[2].copy()

    @computed_property
    def bin_counts(self):
        return self._reduce()[0].copy()

    @computed_property
    def box(self):
        import freud_amd.box

        return freud_amd.box.Box.from_box(self._box)

    @property
    def bin_edges(self):
        """Histogram.h:126-138: min + float(i) * width."""
        width = (self._r_max32 - self._r_min32) / np.float32(self._bins)
        return (self._r_min32 + np.arange(self._bins + 1, dtype=np.float32) * width).astype(np.float32)

    @property
    def bin_centers(self):
        e = self.bin_edges
        return ((e[:-1] + e[1:]) / np.float32(2.0)).astype(np.float32)

    @property
    def bounds(self):
        return (float(self._r_min32), float(self._r_max32))

    @property
    def nbins(self):
        return self._bins

    def __repr__(self):
        return f"freud_amd.density.RDF(bins={self._bins}, r_max={self.bounds[1]}, r_min={self.bounds[0]})"
