"""Connected-component clustering over bonds: the API of ``freud.cluster.Cluster``
(freud/cluster.py, freud/cluster/Cluster.cc) on the GPU."""

import ctypes

import numpy as np

from freud_amd import _lib
from freud_amd._lib import check, lib
from freud_amd.locality import _PairCompute
from freud_amd.util import _Compute, computed_property


class Cluster(_PairCompute):
    """Clusters of points connected by bonds; labels ordered by cluster size
    (descending) then smallest member index (Cluster.cc:110-128)."""

    def __init__(self):
        self._called_compute = False

    def compute(self, system, keys=None, neighbors=None):
        nq, nlist, qargs, d_q, nqp = self._preprocess_arguments(system, neighbors=neighbors)
        nlist = self._neighbor_list(nq, nlist, qargs, d_q, nqp)
        n = nq._n
        self._d_cidx = _lib.reuse(self, "_d_cidx", (n,), np.uint32)
        self._d_sizes = _lib.reuse(self, "_d_sizes", (n,), np.uint32)
        ncl = ctypes.c_uint32()
        check(lib().fa_cluster(n, ctypes.byref(nlist._c), None, self._d_cidx.ptr, self._d_sizes.ptr,
                               ctypes.byref(ncl)))
        self._num_clusters = int(ncl.value)
        self._keys = None if keys is None else np.asarray(keys)
        return self

    @computed_property
    def num_clusters(self):
        return self._num_clusters

    @computed_property
    def cluster_idx(self):
        return self._d_cidx.numpy()

    @computed_property
    def cluster_keys(self):
        idx = self._d_cidx.numpy()
        keys = np.arange(len(idx)) if self._keys is None else self._keys
        order = np.argsort(idx, kind="stable")
        bounds = np.searchsorted(idx[order], np.arange(self._num_clusters + 1))
        return [list(keys[order[bounds[c]: bounds[c + 1]]]) for c in range(self._num_clusters)]

    def __repr__(self):
        return f"freud_amd.cluster.{type(self).__name__}()"
