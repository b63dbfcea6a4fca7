"""Bond-orientational order: the APIs of ``freud.order.Steinhardt``,
``freud.order.SolidLiquid`` and ``freud.order.Hexatic`` (freud/order.py) on the GPU."""

import collections.abc
import ctypes
import warnings
from math import floor

import numpy as np

from freud_amd import _lib
from freud_amd._lib import check, lib
from freud_amd.locality import _PairCompute
from freud_amd.util import _Compute, computed_property


class Steinhardt(_PairCompute):
    """q_l / w_l order parameters (freud/order.py:375-642, Steinhardt.cc)."""

    def __init__(self, l, average=False, wl=False, weighted=False, wl_normalize=False):
        if not isinstance(l, collections.abc.Sequence):
            l = [l]
        if len(l) == 0:
            msg = "At least one l must be specified."
            raise ValueError(msg)
        self._ls = [int(x) for x in l]
        self._average = bool(average)
        self._wl = bool(wl)
        self._weighted = bool(weighted)
        self._wl_normalize = bool(wl_normalize)
        self._called_compute = False

    @property
    def average(self):
        return self._average

    @property
    def wl(self):
        return self._wl

    @property
    def weighted(self):
        return self._weighted

    @property
    def wl_normalize(self):
        return self._wl_normalize

    @property
    def l(self):
        return self._ls[0] if len(self._ls) == 1 else list(self._ls)

    def _flags(self):
        return ((_lib.ST_AVERAGE if self._average else 0) | (_lib.ST_WL if self._wl else 0)
                | (_lib.ST_WEIGHTED if self._weighted else 0) | (_lib.ST_WL_NORMALIZE if self._wl_normalize else 0))

    def compute(self, system, neighbors=None):
        nq, nlist, qargs, d_q, nqp = self._preprocess_arguments(system, neighbors=neighbors)
        nlist = self._neighbor_list(nq, nlist, qargs, d_q, nqp)
        self._run(nq, nlist)
        return self

    def _run(self, nq, nlist):
        n = nq._n
        nl = len(self._ls)
        total_m = sum(2 * l + 1 for l in self._ls)
        self._d_qlm = _lib.reuse(self, "_d_qlm", (n * total_m * 2,), np.float32)
        self._d_ql = _lib.reuse(self, "_d_ql", (n, nl), np.float32)
        self._d_qlave = _lib.reuse(self, "_d_qlave", (n, nl), np.float32) if self._average else None
        self._d_qlmave = _lib.reuse(self, "_d_qlmave", (n * total_m * 2,), np.float32) if self._average else None
        self._d_wl = _lib.reuse(self, "_d_wl", (n, nl), np.float32) if self._wl else None
        order = np.zeros(nl, np.float32)
        ls = np.asarray(self._ls, dtype=np.uint32)
        check(lib().fa_steinhardt(ctypes.byref(nq._cbox), nq._d_points.ptr, n, ctypes.byref(nlist._c),
                                  ls.ctypes.data, nl, self._flags(), self._d_qlm.ptr, self._d_ql.ptr,
                                  self._d_qlave.ptr if self._d_qlave else None,
                                  self._d_qlmave.ptr if self._d_qlmave else None,
                                  self._d_wl.ptr if self._d_wl else None, order.ctypes.data))
        self._order = order
        self._n = n
        self._cache = {}

    def _get(self, name):
        if name not in self._cache:
            if name == "ql":
                self._cache[name] = self._d_ql.numpy()
            elif name == "qlave":
                self._cache[name] = self._d_qlave.numpy()
            elif name == "wl":
                self._cache[name] = self._d_wl.numpy()
            elif name == "qlm":
                flat = self._d_qlm.numpy()
                out, off = [], 0
                for l in self._ls:
                    nm = 2 * l + 1
                    blk = flat[2 * self._n * off: 2 * self._n * (off + nm)].reshape(self._n, nm, 2)
                    out.append((blk[..., 0] + 1j * blk[..., 1]).astype(np.complex64))
                    off += nm
                self._cache[name] = out
        return self._cache[name]

    @computed_property
    def order(self):
        return float(self._order[0]) if len(self._order) == 1 else self._order.copy()

    @computed_property
    def particle_order(self):
        if self._wl:
            p = self._get("wl")
        elif self._average:
            p = self._get("qlave")
        else:
            p = self._get("ql")
        return np.ravel(p).copy() if p.shape[1] == 1 else p.copy()

    @computed_property
    def ql(self):
        q = self._get("ql")
        return np.ravel(q).copy() if q.shape[1] == 1 else q.copy()

    @computed_property
    def particle_harmonics(self):
        q = self._get("qlm")
        return q if len(q) > 1 else q[0]

    def __repr__(self):
        return (
            f"freud_amd.order.{type(self).__name__}(l={self.l}, average={self.average}, wl={self.wl}, "
            f"weighted={self.weighted}, wl_normalize={self.wl_normalize})"
        )


class SolidLiquid(_PairCompute):
    """Solid-like clusters from q_lm bond dot products (freud/order.py:645-813,
    SolidLiquid.cc)."""

    def __init__(self, l, q_threshold, solid_threshold, normalize_q=True):
        if not isinstance(solid_threshold, int):
            warnings.warn(
                f"solid_threshold should be an integer, and will be rounded down (got {solid_threshold})",
                RuntimeWarning,
                stacklevel=2,
            )
            solid_threshold = floor(solid_threshold)
        # SolidLiquid.cc:20-29
        if q_threshold < 0.0:
            msg = "SolidLiquid requires that the dot product cutoff q_threshold must be non-negative."
            raise ValueError(msg)
        self._l = int(l)
        self._q_threshold = float(np.float32(q_threshold))
        self._solid_threshold = int(solid_threshold)
        self._normalize_q = bool(normalize_q)
        self._called_compute = False

    def compute(self, system, neighbors=None):
        nq, nlist, qargs, d_q, nqp = self._preprocess_arguments(system, neighbors=neighbors)
        nlist = self._neighbor_list(nq, nlist, qargs, d_q, nqp)
        n = nq._n
        nm = 2 * self._l + 1
        self._d_qlm = _lib.reuse(self, "_d_qlm", (n, nm, 2), np.float32)
        self._d_ql = _lib.reuse(self, "_d_ql", (n,), np.float32)
        self._d_qlij = _lib.reuse(self, "_d_qlij", (max(len(nlist), 1),), np.float32)
        self._d_nc = _lib.reuse(self, "_d_nc", (n,), np.uint32)
        self._d_cidx = _lib.reuse(self, "_d_cidx", (n,), np.uint32)
        self._d_sizes = _lib.reuse(self, "_d_sizes", (n,), np.uint32)
        ncl = ctypes.c_uint32()
        check(lib().fa_solidliquid(ctypes.byref(nq._cbox), nq._d_points.ptr, n, ctypes.byref(nlist._c), self._l,
                                   self._q_threshold, self._solid_threshold, int(self._normalize_q),
                                   self._d_qlm.ptr, self._d_ql.ptr, self._d_qlij.ptr, self._d_nc.ptr,
                                   self._d_cidx.ptr, self._d_sizes.ptr, ctypes.byref(ncl)))
        self._num_clusters = int(ncl.value)
        self._nlist = nlist
        self._nb = len(nlist)
        return self

    @property
    def l(self):
        return self._l

    @property
    def q_threshold(self):
        return self._q_threshold

    @property
    def solid_threshold(self):
        return self._solid_threshold

    @property
    def normalize_q(self):
        return self._normalize_q

    @computed_property
    def cluster_idx(self):
        return self._d_cidx.numpy()

    @computed_property
    def ql_ij(self):
        return self._d_qlij.numpy()[: self._nb]

    @computed_property
    def particle_harmonics(self):
        q = self._d_qlm.numpy()
        return (q[..., 0] + 1j * q[..., 1]).astype(np.complex64)

    @computed_property
    def cluster_sizes(self):
        return self._d_sizes.numpy()[: self._num_clusters]

    @computed_property
    def largest_cluster_size(self):
        s = self.cluster_sizes
        return int(s[0]) if len(s) else 0

    @computed_property
    def nlist(self):
        return self._nlist

    @computed_property
    def num_connections(self):
        return self._d_nc.numpy()

    def __repr__(self):
        return (
            f"freud_amd.order.{type(self).__name__}(l={self.l}, q_threshold={self.q_threshold}, "
            f"solid_threshold={self.solid_threshold}, normalize_q={self.normalize_q})"
        )


class Hexatic(_PairCompute):
    """k-atic order psi_k for 2D systems (freud/order.py:230-372,
    HexaticTranslational.cc)."""

    def __init__(self, k=6, weighted=False):
        self._k = int(k)
        self._weighted = bool(weighted)
        self._called_compute = False

    def compute(self, system, neighbors=None):
        nq, nlist, qargs, d_q, nqp = self._preprocess_arguments(system, neighbors=neighbors)
        if not nq._box.is2D:
            msg = "A 3D box was provided to a class that only supports 2D systems."
            raise ValueError(msg)
        nlist = self._neighbor_list(nq, nlist, qargs, d_q, nqp)
        self._d_psi = _lib.reuse(self, "_d_psi", (nq._n, 2), np.float32)
        check(lib().fa_hexatic(ctypes.byref(nq._cbox), nq._d_points.ptr, nq._n, ctypes.byref(nlist._c), self._k,
                               int(self._weighted), self._d_psi.ptr))
        return self

    @property
    def default_query_args(self):
        return dict(mode="nearest", num_neighbors=self._k)

    @computed_property
    def particle_order(self):
        p = self._d_psi.numpy()
        return (p[:, 0] + 1j * p[:, 1]).astype(np.complex64)

    @property
    def k(self):
        return self._k

    @property
    def weighted(self):
        return self._weighted

    def __repr__(self):
        return f"freud_amd.order.{type(self).__name__}(k={self.k}, weighted={self.weighted})"
