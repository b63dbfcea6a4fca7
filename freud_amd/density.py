"""Radial distribution function, local density and the pair correlation of
per-point values: the APIs of ``freud.density.RDF`` (freud/density.py:535-696),
``freud.density.LocalDensity`` (freud/density.py:418-515) and
``freud.density.CorrelationFunction`` (freud/density.py:28-190) on the GPU.

Without a NeighborList the bins are filled by the fused query + histogram
kernel (``fa_rdf_accumulate_query``); a given NeighborList is binned from its
distances.  Counts stay on the device across frames (``reset=False``), and the
normalisation (RDF.cc:73-99) runs in the reference's float order.
"""

import ctypes

import numpy as np

from freud_amd import _lib
from freud_amd._lib import DeviceArray, check, lib
from freud_amd.locality import _PairCompute, _SpatialHistogram1D
from freud_amd.util import _Compute, _convert_array, computed_property


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


class LocalDensity(_PairCompute):
    """Local number density around each query point (freud/density.py:418-515,
    freud/density/LocalDensity.cc): points within r_max - diameter/2 count
    fully, points whose centres lie within diameter/2 of r_max count linearly
    in between.  Sums run in NeighborList bond order on the device."""

    def __init__(self, r_max, diameter):
        # LocalDensity.cc:26-37
        if r_max <= 0:
            msg = "LocalDensity requires r_max to be positive."
            raise ValueError(msg)
        if diameter < 0:
            msg = "LocalDensity requires diameter to be non-negative."
            raise ValueError(msg)
        self._r_max32 = np.float32(r_max)
        self._diameter32 = np.float32(diameter)

    @property
    def r_max(self):
        return float(self._r_max32)

    @property
    def diameter(self):
        return float(self._diameter32)

    @property
    def default_query_args(self):
        """``{'mode': 'ball', 'r_max': r_max + 0.5 * diameter}`` (density.py:510-515)."""
        return dict(mode="ball", r_max=self.r_max + 0.5 * self.diameter)

    def compute(self, system, query_points=None, neighbors=None):
        nq, nlist, qargs, d_q, num_query_points = self._preprocess_arguments(system, query_points, neighbors)
        nlist = self._neighbor_list(nq, nlist, qargs, d_q, num_query_points)
        n = int(num_query_points)
        dens = _lib.reuse(self, "_d_density", (max(n, 1),), np.float32)
        num = _lib.reuse(self, "_d_num", (max(n, 1),), np.float32)
        check(lib().fa_local_density(ctypes.byref(nlist._c), float(self._r_max32), float(self._diameter32),
                                     int(bool(nq._box.is2D)), dens.ptr, num.ptr))
        self._n = n
        self._box = nq._box
        self._host = None
        return self

    def _results(self):
        if self._host is None:
            self._host = (self._d_density.numpy()[: self._n], self._d_num.numpy()[: self._n])
        return self._host

    @computed_property
    def density(self):
        return self._results()[0].copy()

    @computed_property
    def num_neighbors(self):
        return self._results()[1].copy()

    @computed_property
    def box(self):
        import freud_amd.box

        return freud_amd.box.Box.from_box(self._box)

    def __repr__(self):
        return f"freud_amd.density.LocalDensity(r_max={self.r_max}, diameter={self.diameter})"


class CorrelationFunction(_SpatialHistogram1D):
    """Average product conj(s_1(point)) s_2(query point) of per-point values
    over bonds at distance r (freud/density.py:28-190,
    freud/density/CorrelationFunction.cc).  Counts and complex double sums
    accumulate on the device across frames (``reset=False``)."""

    def __init__(self, bins, r_max):
        # CorrelationFunction.cc:27-35
        if bins == 0:
            msg = "CorrelationFunction  requires a nonzero number of bins."
            raise ValueError(msg)
        if r_max <= 0:
            msg = "CorrelationFunction requires r_max to be positive."
            raise ValueError(msg)
        self._bins = int(bins)
        self._r_max32 = np.float32(r_max)
        self.r_max = r_max
        self.is_complex = False
        self._counts = None
        self._sums = None

    def _reset(self):
        if self._counts is None:
            self._counts = DeviceArray.zeros(self._bins, np.uint32)
            self._sums = DeviceArray.zeros(2 * self._bins, np.float64)
        else:
            self._counts.zero()
            self._sums.zero()
        self._reduced = None

    def compute(self, system, values, query_points=None, query_values=None, neighbors=None, reset=True):
        if reset or self._counts is None:
            self.is_complex = False
            self._reset()
        nq, nlist, qargs, d_q, num_query_points = self._preprocess_arguments(system, query_points, neighbors)
        # freud/density.py:109-128
        self.is_complex = bool(self.is_complex or np.any(np.iscomplex(values)) or np.any(np.iscomplex(query_values)))
        values = _convert_array(values, shape=(nq._n,), dtype=np.complex128)
        if query_values is None:
            query_values = values
        else:
            query_values = _convert_array(query_values, shape=(num_query_points,), dtype=np.complex128)
        nlist = self._neighbor_list(nq, nlist, qargs, d_q, num_query_points)
        check(lib().fa_correlation_accumulate(ctypes.byref(nlist._c), values.ctypes.data, query_values.ctypes.data,
                                              self._bins, float(self._r_max32), self._counts.ptr, self._sums.ptr))
        self._box = nq._box
        self._reduced = None
        return self

    def _reduce(self):
        if self._reduced is None:
            counts = self._counts.numpy()
            sums = self._sums.numpy()
            out = np.empty(2 * self._bins, np.float64)
            check(lib().fa_correlation_reduce(counts.ctypes.data, sums.ctypes.data, self._bins, out.ctypes.data))
            self._reduced = (counts, out.view(np.complex128))
        return self._reduced

    @computed_property
    def correlation(self):
        out = self._reduce()[1].copy()
        return out if self.is_complex else np.real(out)

    @computed_property
    def bin_counts(self):
        return self._reduce()[0].copy()

    @computed_property
    def box(self):
        import freud_amd.box

        return freud_amd.box.Box.from_box(self._box)

    @property
    def bin_edges(self):
        """RegularAxis(bins, 0, r_max) edges (Histogram.h:126-138)."""
        width = (self._r_max32 - np.float32(0)) / np.float32(self._bins)
        return (np.float32(0) + np.arange(self._bins + 1, dtype=np.float32) * width).astype(np.float32)

    @property
    def bin_centers(self):
        e = self.bin_edges
        return ((e[:-1] + e[1:]) / np.float32(2.0)).astype(np.float32)

    @property
    def bounds(self):
        return (0.0, float(self._r_max32))

    @property
    def nbins(self):
        return self._bins

    @property
    def bins(self):
        return self._bins

    def __repr__(self):
        return f"freud_amd.density.CorrelationFunction(bins={self.nbins}, r_max={self.r_max})"
