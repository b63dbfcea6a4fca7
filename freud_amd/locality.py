"""Neighbour finding: the API of ``freud.locality`` (freud/locality.py) backed by
the HIP kernels of ``freud_amd/csrc`` through the C-ABI of ``include/freud_amd.h``.

Points and NeighborLists live in device memory; numpy views of NeighborList
arrays are copied from the device the first time they are read.
"""

import ctypes
import inspect

import numpy as np

import freud_amd.box
from freud_amd import _lib
from freud_amd._lib import DeviceArray, check, lib
from freud_amd.util import NO_DEFAULT_QUERY_ARGS_MESSAGE, _Compute, _convert_array, _convert_box, computed_property

_DEFAULT_K = 0xFFFFFFFF


class _QueryArgs:
    """Query arguments (freud/locality.py:39-175, NeighborQuery.h:37-73)."""

    def __init__(self, mode=None, r_min=None, r_max=None, r_guess=None, num_neighbors=None, exclude_ii=None,
                 scale=None, **kwargs):
        self._mode = 0
        self._r_min = 0.0
        self._r_max = -1.0
        self._r_guess = -1.0
        self._num_neighbors = _DEFAULT_K
        self._exclude_ii = False
        self._scale = -1.0
        self.mode = mode
        if r_max is not None:
            self.r_max = r_max
        if r_min is not None:
            self.r_min = r_min
        if r_guess is not None:
            self.r_guess = r_guess
        if num_neighbors is not None:
            self.num_neighbors = num_neighbors
        if exclude_ii is not None:
            self.exclude_ii = exclude_ii
        if scale is not None:
            self.scale = scale
        if len(kwargs) > 0:
            err_str = ", ".join(f"{k} = {v}" for k, v in kwargs.items())
            raise ValueError("The following invalid query arguments were provided: " + err_str)

    def update(self, qargs):
        if qargs is None:
            return
        for arg in qargs:
            if hasattr(self, arg):
                setattr(self, arg, qargs[arg])
            else:
                msg = "An invalid query argument was provided."
                raise ValueError(msg)

    @classmethod
    def from_dict(cls, mapping):
        return cls(**mapping)

    @property
    def mode(self):
        return {0: None, 1: "ball", 2: "nearest"}[self._mode]

    @mode.setter
    def mode(self, value):
        if value == "none" or value is None:
            self._mode = 0
        elif value == "ball":
            self._mode = 1
        elif value == "nearest":
            self._mode = 2
        else:
            msg = "An invalid mode was provided."
            raise ValueError(msg)

    @property
    def r_guess(self):
        return self._r_guess

    @r_guess.setter
    def r_guess(self, value):
        self._r_guess = float(np.float32(value))

    @property
    def r_min(self):
        return self._r_min

    @r_min.setter
    def r_min(self, value):
        self._r_min = float(np.float32(value))

    @property
    def r_max(self):
        return self._r_max

    @r_max.setter
    def r_max(self, value):
        self._r_max = float(np.float32(value))

    @property
    def num_neighbors(self):
        return self._num_neighbors

    @num_neighbors.setter
    def num_neighbors(self, value):
        self._num_neighbors = int(value) & 0xFFFFFFFF

    @property
    def exclude_ii(self):
        return self._exclude_ii

    @exclude_ii.setter
    def exclude_ii(self, value):
        self._exclude_ii = bool(value)

    @property
    def scale(self):
        return self._scale

    @scale.setter
    def scale(self, value):
        self._scale = float(np.float32(value))

    def _c(self):
        a = _lib.fa_query_args()
        a.mode = self._mode
        a.num_neighbors = self._num_neighbors
        a.r_max = self._r_max
        a.r_min = self._r_min
        a.r_guess = self._r_guess
        a.scale = self._scale
        a.exclude_ii = int(self._exclude_ii)
        return a

    def __repr__(self):
        return (
            f"freud_amd.locality.{type(self).__name__}(mode={self.mode}, r_max={self.r_max}, "
            f"num_neighbors={self.num_neighbors}, exclude_ii={self.exclude_ii}, scale={self.scale})"
        )


class NeighborQueryResult:
    """Lazy query result (freud/locality.py:178-229)."""

    def __init__(self, nq, points, query_args):
        self._nq = nq
        self._points = np.array(points, dtype=np.float32)
        self._query_args = query_args

    def __iter__(self):
        nl = self.toNeighborList()
        for (i, j), d in zip(nl[:], nl.distances):
            yield (int(i), int(j), float(d))

    def toNeighborList(self, sort_by_distance=False):
        q = self._points.reshape(-1, 3)
        host = self._nq._host_points
        if host is not None and host.shape == q.shape and np.array_equal(host.view(np.uint32), q.view(np.uint32)):
            # the query points are the NeighborQuery's own points: query with its
            # device copy (no upload; the library then runs its self-query kernels)
            d_q = self._nq._d_points
        else:
            d_q = DeviceArray.from_numpy(q)
        return self._nq._run_query(d_q, len(q), self._query_args, sort_by_distance)


class NeighborQuery:
    """Points plus a box that can be queried for neighbours (freud/locality.py:232-440)."""

    _kind = _lib.NQ_AABB

    def __init__(self):
        msg = "The NeighborQuery class is abstract, and should not be directly instantiated"
        raise RuntimeError(msg)

    def _setup(self, box, points, cell_width=0.0):
        b = _convert_box(box)
        self._box = b
        if isinstance(points, DeviceArray):
            if points.dtype != np.float32 or len(points.shape) != 2 or points.shape[1] != 3:
                msg = "device points must be float32 of shape (N, 3)"
                raise ValueError(msg)
            self._d_points = points
            self._host_points = None
            n = points.shape[0]
        else:
            pts = _convert_array(points, shape=(None, 3)).copy()
            # NeighborQuery.h:103-112
            if b.is2D and len(pts) and np.any(np.abs(pts[:, 2]) > 1e-6):
                msg = "A point with z != 0 was provided in a 2D box."
                raise ValueError(msg)
            self._host_points = pts
            n = len(pts)
            if n == 0:
                msg = "Cannot create a NeighborQuery with 0 particles."
                raise ValueError(msg)
            self._d_points = DeviceArray.from_numpy(pts)
        self._n = n
        self._cbox = _lib.make_box(b)
        h = ctypes.c_void_p()
        check(lib().fa_nq_create(ctypes.byref(self._cbox), self._d_points.ptr, n, self._kind, float(cell_width),
                                 ctypes.byref(h)))
        self._handle = h.value

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h and _lib._lib is not None:
            _lib._lib.fa_nq_destroy(h)
            self._handle = None

    @classmethod
    def from_system(cls, system, dimensions=None):
        """freud/locality.py:268-382."""

        def _match_class_path(obj, *matches):
            return any(c.__module__ + "." + c.__name__ in matches for c in inspect.getmro(type(obj)))

        if isinstance(system, cls):
            return system
        if isinstance(system, NeighborQuery) and cls is NeighborQuery:
            return system
        if _match_class_path(system, "gsd.hoomd.Frame", "gsd.hoomd.Snapshot", "hoomd.snapshot.Snapshot"):
            box = np.array(system.configuration.box)
            if system.configuration.dimensions == 2:
                box[[2, 4, 5]] = 0
            system = (box, system.particles.position)
        elif hasattr(system, "box") and hasattr(system, "particles") and hasattr(system.particles, "position"):
            system = (system.box, system.particles.position)
        elif hasattr(system, "box") and hasattr(system, "points"):
            box = system.box
            if hasattr(system, "freud_box"):
                box = system.freud_box()
            system = (box, system.points)
        if cls == NeighborQuery:
            return _RawPoints(*system)
        return cls(*system)

    @property
    def box(self):
        return freud_amd.box.Box.from_box(self._box)

    @property
    def points(self):
        if self._host_points is None:
            self._host_points = self._d_points.numpy()
        return np.asarray(self._host_points)

    def query(self, query_points, query_args):
        query_points = _convert_array(np.atleast_2d(query_points), shape=(None, 3))
        args = _QueryArgs.from_dict(query_args)
        return NeighborQueryResult(self, query_points, args)

    # run a query with device query points, returning a NeighborList
    def _run_query(self, d_query_points, n_query, qargs, sort_by_distance=False):
        nl = NeighborList()
        a = qargs._c()
        check(lib().fa_query(self._handle, d_query_points.ptr, int(n_query), ctypes.byref(a),
                             int(bool(sort_by_distance)), ctypes.byref(nl._c)))
        nl._owner_refs = (self, d_query_points)
        return nl


class NeighborList:
    """Bonds between query points and points (freud/locality.py:443-751),
    stored on the device in freud's layout (NeighborList.h:29-158)."""

    def __init__(self, _null=False):
        self._c = _lib.fa_nlist()
        self._cache = {}
        self._owner_refs = None
        self._null = _null

    def __del__(self):
        c = getattr(self, "_c", None)
        if c is not None and _lib._lib is not None and c.neighbors:
            _lib._lib.fa_nlist_free(ctypes.byref(c))

    def _invalidate(self):
        self._cache = {}

    def _host(self, name):
        if name not in self._cache:
            nb = int(self._c.num_bonds)
            nq = int(self._c.num_query_points)
            if name == "neighbors":
                a = _lib.copy_d2h(self._c.neighbors, (nb, 2), np.uint32) if nb else np.zeros((0, 2), np.uint32)
            elif name == "distances":
                a = _lib.copy_d2h(self._c.distances, (nb,), np.float32) if nb else np.zeros(0, np.float32)
            elif name == "weights":
                a = _lib.copy_d2h(self._c.weights, (nb,), np.float32) if nb else np.zeros(0, np.float32)
            elif name == "vectors":
                a = _lib.copy_d2h(self._c.vectors, (nb, 3), np.float32) if nb else np.zeros((0, 3), np.float32)
            elif name == "counts":
                a = _lib.copy_d2h(self._c.counts, (nq,), np.uint32) if nq and self._c.counts else np.zeros(
                    nq, np.uint32)
            elif name == "segments":
                a = _lib.copy_d2h(self._c.segments, (nq,), np.uint32) if nq and self._c.segments else np.zeros(
                    nq, np.uint32)
            else:
                raise KeyError(name)
            a.flags.writeable = False
            self._cache[name] = a
        return self._cache[name]

    @classmethod
    def from_arrays(cls, num_query_points, num_points, query_point_indices, point_indices, vectors, weights=None):
        """freud/locality.py:480-560 and NeighborList.cc:51-82."""
        query_point_indices = _convert_array(query_point_indices, shape=(None,), dtype=np.uint32).copy()
        point_indices = _convert_array(point_indices, shape=query_point_indices.shape, dtype=np.uint32).copy()
        vectors = _convert_array(vectors, shape=(len(query_point_indices), 3), dtype=np.float32).copy()
        if weights is None:
            weights = np.ones(query_point_indices.shape, dtype=np.float32)
        weights = _convert_array(weights, shape=query_point_indices.shape).copy()
        result = cls()
        dq = DeviceArray.from_numpy(query_point_indices)
        dp = DeviceArray.from_numpy(point_indices)
        dv = DeviceArray.from_numpy(vectors)
        dw = DeviceArray.from_numpy(weights)
        check(lib().fa_nlist_from_arrays(int(num_query_points), int(num_points), dq.ptr, dp.ptr, dv.ptr, dw.ptr,
                                         len(query_point_indices), ctypes.byref(result._c)))
        return result

    @classmethod
    def all_pairs(cls, system, query_points=None, exclude_ii=True):
        """freud/locality.py:562-600, NeighborList.cc:84-128."""
        nq = NeighborQuery.from_system(system)
        points = nq._d_points
        n_points = nq._n
        if query_points is None:
            d_q = points
            n_q = n_points
        else:
            qp = _convert_array(query_points, shape=(None, 3))
            d_q = DeviceArray.from_numpy(qp)
            n_q = len(qp)
        result = cls()
        check(lib().fa_nlist_all_pairs(ctypes.byref(nq._cbox), points.ptr, n_points, d_q.ptr, n_q,
                                       int(bool(exclude_ii)), ctypes.byref(result._c)))
        return result

    def copy(self, other=None):
        if other is not None:
            assert isinstance(other, NeighborList)
            check(lib().fa_nlist_copy(ctypes.byref(other._c), ctypes.byref(self._c)))
            self._invalidate()
            return self
        new_copy = NeighborList()
        new_copy.copy(self)
        return new_copy

    def __getitem__(self, key):
        return self._host("neighbors")[key]

    @property
    def query_point_indices(self):
        return self[:, 0]

    @property
    def point_indices(self):
        return self[:, 1]

    @property
    def weights(self):
        return self._host("weights")

    @property
    def distances(self):
        return self._host("distances")

    @property
    def vectors(self):
        return self._host("vectors")

    @property
    def segments(self):
        return self._host("segments")

    @property
    def neighbor_counts(self):
        return self._host("counts")

    def __len__(self):
        return int(self._c.num_bonds)

    @property
    def num_query_points(self):
        return int(self._c.num_query_points)

    @property
    def num_points(self):
        return int(self._c.num_points)

    def find_first_index(self, i):
        """NeighborList.cc:309-316, 416-430 (bisection over query indices)."""
        nb = len(self)
        if nb == 0:
            return 0
        q = self.query_point_indices

        left, right = 0, nb
        while left + 1 < right:
            middle = (left + right) // 2
            if q[middle] < i:
                left = middle
            else:
                right = middle
        return left + (1 if i > q[0] else 0)

    def filter(self, filt):
        filt = np.ascontiguousarray(filt, dtype=bool)
        if len(filt) < len(self):
            msg = "The filter must have one entry per bond."
            raise ValueError(msg)
        d = DeviceArray.from_numpy(filt.view(np.uint8))
        check(lib().fa_nlist_filter(ctypes.byref(self._c), d.ptr))
        self._invalidate()
        return self

    def filter_r(self, r_max, r_min=0):
        check(lib().fa_nlist_filter_r(ctypes.byref(self._c), float(r_max), float(r_min)))
        self._invalidate()
        return self

    def sort(self, by_distance=False):
        check(lib().fa_nlist_sort(ctypes.byref(self._c), int(bool(by_distance))))
        self._invalidate()
        return self

    def _validate(self, num_query_points, num_points):
        # NeighborList.cc:348-358
        if num_query_points != self.num_query_points or num_points != self.num_points:
            msg = "NeighborList found inconsistent array sizes."
            raise RuntimeError(msg)


def _make_default_nq(neighbor_query):
    if not isinstance(neighbor_query, NeighborQuery):
        return _RawPoints(*neighbor_query)
    return neighbor_query


class _RawPoints(NeighborQuery):
    """Box + points; queries go through the AABBQuery arithmetic (RawPoints.h:50-72)."""

    _kind = _lib.NQ_AABB

    def __init__(self, box, points):
        self._setup(box, points)


class AABBQuery(NeighborQuery):
    """freud.locality.AABBQuery (freud/locality.py:852-869): bond vectors
    r = p_j - (q_i + image) (AABBQuery.cc:93,125)."""

    _kind = _lib.NQ_AABB

    def __init__(self, box, points):
        self._setup(box, points)


class LinkCell(NeighborQuery):
    """freud.locality.LinkCell (freud/locality.py:872-897): bond vectors
    r = box.wrap(p_j - q_i) (LinkCell.cc:522,616)."""

    _kind = _lib.NQ_LINKCELL

    def __init__(self, box, points, cell_width=0):
        self._setup(box, points, cell_width)

    @property
    def cell_width(self):
        w = ctypes.c_float()
        check(lib().fa_nq_cell_width(self._handle, ctypes.byref(w)))
        return w.value


class CellQuery(NeighborQuery):
    """freud.locality.CellQuery (freud/locality.py:900-921): ball queries over
    a cell grid of width r_max holding the points and their ghost copies near
    the faces; bond vectors r = (p_j + ghost shift) - q_i
    (CellIterator.h:177-186).  Nearest-neighbour queries raise RuntimeError
    (CellQuery.h:178-183)."""

    _kind = _lib.NQ_CELLQUERY

    def __init__(self, box, points):
        self._setup(box, points)


class _PairCompute(_Compute):
    """freud/locality.py:924-1021."""

    def _preprocess_arguments(self, system, query_points=None, neighbors=None):
        nq = NeighborQuery.from_system(system)
        nlist, qargs = self._resolve_neighbors(neighbors, query_points)
        if query_points is None:
            d_q = nq._d_points
            num_query_points = nq._n
        else:
            query_points = _convert_array(query_points, shape=(None, 3))
            d_q = DeviceArray.from_numpy(query_points)
            num_query_points = query_points.shape[0]
        return (nq, nlist, qargs, d_q, num_query_points)

    def _resolve_neighbors(self, neighbors, query_points=None):
        if type(neighbors) is NeighborList:
            nlist = neighbors
            qargs = _QueryArgs()
        elif neighbors is None or type(neighbors) is dict:
            query_args = self.default_query_args if neighbors is None else neighbors.copy()
            query_args.setdefault("exclude_ii", query_points is None)
            qargs = _QueryArgs.from_dict(query_args)
            nlist = None
        else:
            msg = "An invalid value was provided for neighbors, which must be a dict or NeighborList object."
            raise ValueError(msg)
        return nlist, qargs

    @staticmethod
    def _neighbor_list(nq, nlist, qargs, d_q, num_query_points):
        """The given NeighborList, or one made by querying (makeDefaultNlist,
        NeighborComputeFunctional.cc:16-33)."""
        if nlist is None:
            nlist = nq._run_query(d_q, num_query_points, qargs)
        nlist._validate(num_query_points, nq._n)
        return nlist

    @property
    def default_query_args(self):
        raise NotImplementedError(NO_DEFAULT_QUERY_ARGS_MESSAGE.format(type(self).__name__))


class _SpatialHistogram(_PairCompute):
    @property
    def default_query_args(self):
        return dict(mode="ball", r_max=self.r_max)


class _SpatialHistogram1D(_SpatialHistogram):
    pass


class Filter(_PairCompute):
    """Filter an existing NeighborList (freud/locality.py:1271-1336).

    ``compute`` builds the unfiltered list from ``neighbors`` (a NeighborList,
    query arguments, or every pair when None) and keeps the subset of its bonds
    the derived filter selects; ``filtered_nlist`` is sorted by distance.  The
    filtering runs on the device (fa_filter_sann / fa_filter_rad).
    """

    def __init__(self):
        msg = "The Filter class is abstract and should not be instantiated directly."
        raise RuntimeError(msg)

    def _setup_filter(self, allow_incomplete_shell):
        _Compute.__init__(self)
        self._allow_incomplete_shell = bool(allow_incomplete_shell)
        self._filtered = None
        self._unfiltered = None

    def _preprocess_arguments(self, system, query_points=None, neighbors=None):
        """Every pair (without i == i for self queries) when neighbors is None."""
        nq = NeighborQuery.from_system(system)
        if neighbors is None:
            neighbors = NeighborList.all_pairs(nq, query_points, query_points is None)
        return super()._preprocess_arguments(nq, query_points, neighbors)

    def compute(self, system, neighbors=None, query_points=None):
        nq, nlist, qargs, d_q, num_query_points = self._preprocess_arguments(system, query_points, neighbors)
        nlist = self._neighbor_list(nq, nlist, qargs, d_q, num_query_points)
        self._filter(nq, nlist, d_q, None)
        return self

    def _run_filter(self, nlist, call):
        """call(out, d_unfilled, n_unfilled) runs the C-ABI filter; then the
        unfilled shells are reported as Filter.h:80-102 does."""
        out = NeighborList()
        nqp = nlist.num_query_points
        d_unfilled = DeviceArray((max(nqp, 1),), np.uint8)
        n_unfilled = ctypes.c_uint32()
        check(call(out, d_unfilled, n_unfilled))
        self._unfiltered = nlist
        if n_unfilled.value:
            idx = np.flatnonzero(d_unfilled.numpy()[:nqp])
            msg = "Query point indices " + ", ".join(str(int(i)) for i in idx) + " do not have full neighbor shells."
            if not self._allow_incomplete_shell:
                raise RuntimeError(msg)
            print("WARNING: " + msg)
        self._filtered = out

    @computed_property
    def filtered_nlist(self):
        """:class:`NeighborList`: the filtered neighbor list."""
        return self._filtered

    @computed_property
    def unfiltered_nlist(self):
        """:class:`NeighborList`: the unfiltered neighbor list."""
        return self._unfiltered


class FilterSANN(Filter):
    """Solid Angle Nearest Neighbour filter (freud/locality.py:1339-1376,
    FilterSANN.cc:22-104): per query point, the closest bonds whose solid
    angles first sum to 4 pi; kept bonds keep their vectors and weights."""

    def __init__(self, allow_incomplete_shell=False):
        self._setup_filter(allow_incomplete_shell)

    def _filter(self, nq, nlist, d_q, radii):
        self._run_filter(nlist, lambda out, d_u, n_u: lib().fa_filter_sann(
            ctypes.byref(nlist._c), ctypes.byref(out._c), d_u.ptr, ctypes.byref(n_u)))

    def __repr__(self):
        return f"freud_amd.locality.{type(self).__name__}(allow_incomplete_shell={self._allow_incomplete_shell})"


class FilterRAD(Filter):
    """Relative Angular Distance filter (freud/locality.py:1379-1502,
    FilterRAD.cc:22-131): drops each neighbour j blocked by a closer k;
    RAD-closed (terminate_after_blocked) also drops everything past the first
    blocked neighbour.  Kept bonds carry box.wrap(q_i - p_j) and weight 1."""

    def __init__(self, allow_incomplete_shell=False, terminate_after_blocked=True):
        self._setup_filter(allow_incomplete_shell)
        self._terminate_after_blocked = bool(terminate_after_blocked)

    def compute(self, system, neighbors=None, query_points=None, points_radii=None, query_points_radii=None):
        nq, nlist, qargs, d_q, num_query_points = self._preprocess_arguments(system, query_points, neighbors)
        # freud/locality.py:1472-1496
        if points_radii is None:
            points_radii = np.ones(nq._n, dtype=np.float32)
        points_radii = _convert_array(points_radii, shape=(nq._n,))
        if query_points_radii is None:
            query_points_radii = np.ones(num_query_points, dtype=np.float32)
        query_points_radii = _convert_array(query_points_radii, shape=(num_query_points,))
        if not np.all(np.isfinite(points_radii)):
            msg = "points_radii must contain only finite values."
            raise ValueError(msg)
        if not np.all(points_radii > 0):
            msg = "points_radii must contain only values greater than 0."
            raise ValueError(msg)
        if not np.all(np.isfinite(query_points_radii)):
            msg = "query_points_radii must contain only finite values."
            raise ValueError(msg)
        if not np.all(query_points_radii > 0):
            msg = "query_points_radii must contain only values greater than 0."
            raise ValueError(msg)
        nlist = self._neighbor_list(nq, nlist, qargs, d_q, num_query_points)
        # the blocking test reads the radii of points only (FilterRAD.cc:26)
        self._filter(nq, nlist, d_q, points_radii)
        return self

    def _filter(self, nq, nlist, d_q, radii):
        d_r = None if radii is None or np.all(radii == 1) else DeviceArray.from_numpy(radii)
        self._run_filter(nlist, lambda out, d_u, n_u: lib().fa_filter_rad(
            ctypes.byref(nq._cbox), nq._d_points.ptr, nq._n, d_q.ptr, None if d_r is None else d_r.ptr,
            ctypes.byref(nlist._c), int(self._terminate_after_blocked), ctypes.byref(out._c), d_u.ptr,
            ctypes.byref(n_u)))

    def __repr__(self):
        return (f"freud_amd.locality.{type(self).__name__}(allow_incomplete_shell={self._allow_incomplete_shell}, "
                f"terminate_after_blocked={self._terminate_after_blocked})")

