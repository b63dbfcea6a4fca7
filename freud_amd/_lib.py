"""ctypes binding of the HIP library (``freud_amd/_lib/libfreud_amd.so``).

The library is loaded from the package tree; a missing library is an error
(there is no CPU fallback).  Device arrays are wrapped by :class:`DeviceArray`,
which owns a device allocation and copies to and from numpy on request.
"""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libfreud_amd.so")

FA_OK = 0
FA_ERR_INVALID = -1
FA_ERR_RUNTIME = -2
FA_ERR_RANGE = -3
FA_ERR_DEVICE = -4

NQ_AABB = 0
NQ_LINKCELL = 1
NQ_CELLQUERY = 2

PMFT_XYT = 0
PMFT_R12 = 1
PMFT_XYZ = 2

MODE_NONE = 0
MODE_BALL = 1
MODE_NEAREST = 2

ST_AVERAGE = 1
ST_WL = 2
ST_WEIGHTED = 4
ST_WL_NORMALIZE = 8


class fa_box(ctypes.Structure):
    _fields_ = [
        ("Lx", ctypes.c_float),
        ("Ly", ctypes.c_float),
        ("Lz", ctypes.c_float),
        ("xy", ctypes.c_float),
        ("xz", ctypes.c_float),
        ("yz", ctypes.c_float),
        ("is2D", ctypes.c_int32),
    ]


class fa_query_args(ctypes.Structure):
    _fields_ = [
        ("mode", ctypes.c_int32),
        ("num_neighbors", ctypes.c_uint32),
        ("r_max", ctypes.c_float),
        ("r_min", ctypes.c_float),
        ("r_guess", ctypes.c_float),
        ("scale", ctypes.c_float),
        ("exclude_ii", ctypes.c_int32),
    ]


class fa_batch_plan(ctypes.Structure):
    _fields_ = [
        ("ball_r_max", ctypes.c_float),
        ("rdf_bins", ctypes.c_uint32),
        ("rdf_r_max", ctypes.c_float),
        ("l", ctypes.c_uint32),
    ]


class fa_nlist(ctypes.Structure):
    _fields_ = [
        ("neighbors", ctypes.c_void_p),
        ("distances", ctypes.c_void_p),
        ("weights", ctypes.c_void_p),
        ("vectors", ctypes.c_void_p),
        ("counts", ctypes.c_void_p),
        ("segments", ctypes.c_void_p),
        ("offsets", ctypes.c_void_p),
        ("num_bonds", ctypes.c_uint64),
        ("num_query_points", ctypes.c_uint32),
        ("num_points", ctypes.c_uint32),
        ("capacity", ctypes.c_uint64),
    ]


_lib = None

_SIGS = {
    "fa_last_error": (ctypes.c_char_p, []),
    "fa_init": (ctypes.c_int, []),
    "fa_device_name": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int]),
    "fa_malloc": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]),
    "fa_free": (ctypes.c_int, [ctypes.c_void_p]),
    "fa_memcpy_h2d": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]),
    "fa_memcpy_d2h": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]),
    "fa_memcpy_d2d": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]),
    "fa_memset": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]),
    "fa_synchronize": (ctypes.c_int, []),
    "fa_timer_start": (ctypes.c_int, []),
    "fa_timer_stop": (ctypes.c_int, [ctypes.POINTER(ctypes.c_float)]),
    "fa_batch_create": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.POINTER(fa_batch_plan), ctypes.POINTER(ctypes.c_void_p)]),
    "fa_batch_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "fa_batch_points": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]),
    "fa_batch_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "fa_batch_num_bonds": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "fa_batch_nlist": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(fa_nlist)]),
    "fa_batch_rdf_counts": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]),
    "fa_batch_ql": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p),
                                   ctypes.c_void_p]),
    "fa_batch_config": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                       ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_uint64),
                                       ctypes.POINTER(ctypes.c_int)]),
    "fa_sl_features": (ctypes.c_int, [ctypes.POINTER(fa_box), ctypes.c_void_p, ctypes.c_uint32,
                                      ctypes.POINTER(fa_nlist), ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_void_p]),
    "fa_sl_decide": (ctypes.c_int, [ctypes.c_uint32, ctypes.POINTER(fa_nlist), ctypes.c_void_p, ctypes.c_float,
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_uint32,
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32)]),
    "fa_mlp_create": (ctypes.c_int, [ctypes.c_uint64, ctypes.POINTER(ctypes.c_void_p)]),
    "fa_mlp_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "fa_mlp_num_params": (ctypes.c_int, []),
    "fa_mlp_set_params": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "fa_mlp_get_params": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "fa_mlp_set_norm": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "fa_mlp_forward": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]),
    "fa_mlp_gradient": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_uint32, ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]),
    "fa_mlp_train": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                    ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, ctypes.c_uint64,
                                    ctypes.c_void_p]),
    "fa_tune_enable": (ctypes.c_int, [ctypes.c_int]),
    "fa_tune_count": (ctypes.c_int, []),
    "fa_tune_get": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                   ctypes.POINTER(ctypes.c_int)]),
    "fa_tune_set": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "fa_tune_reset": (ctypes.c_int, []),
    "fa_tune_variants": (ctypes.c_int, [ctypes.c_int, ctypes.c_int]),
    "fa_profile_enable": (ctypes.c_int, [ctypes.c_int]),
    "fa_profile_reset": (ctypes.c_int, []),
    "fa_profile_count": (ctypes.c_int, []),
    "fa_profile_get": (ctypes.c_int, [ctypes.c_int, ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64),
                                      ctypes.POINTER(ctypes.c_double)]),
    "fa_box_transform": (ctypes.c_int, [ctypes.POINTER(fa_box), ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_uint64]),
    "fa_nq_create": (ctypes.c_int, [ctypes.POINTER(fa_box), ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int,
                                    ctypes.c_float, ctypes.POINTER(ctypes.c_void_p)]),
    "fa_nq_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "fa_nq_cell_width": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]),
    "fa_cellquery_grid": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.POINTER(ctypes.c_float)]),
    "fa_query": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(fa_query_args),
                                ctypes.c_int, ctypes.POINTER(fa_nlist)]),
    "fa_nlist_free": (ctypes.c_int, [ctypes.POINTER(fa_nlist)]),
    "fa_nlist_from_arrays": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64,
                                            ctypes.POINTER(fa_nlist)]),
    "fa_nlist_all_pairs": (ctypes.c_int, [ctypes.POINTER(fa_box), ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p,
                                          ctypes.c_uint32, ctypes.c_int, ctypes.POINTER(fa_nlist)]),
    "fa_nlist_filter": (ctypes.c_int, [ctypes.POINTER(fa_nlist), ctypes.c_void_p]),
    "fa_nlist_filter_r": (ctypes.c_int, [ctypes.POINTER(fa_nlist), ctypes.c_float, ctypes.c_float]),
    "fa_nlist_sort": (ctypes.c_int, [ctypes.POINTER(fa_nlist), ctypes.c_int]),
    "fa_nlist_copy": (ctypes.c_int, [ctypes.POINTER(fa_nlist), ctypes.POINTER(fa_nlist)]),
    "fa_filter_sann": (ctypes.c_int, [ctypes.POINTER(fa_nlist), ctypes.POINTER(fa_nlist), ctypes.c_void_p,
                                      ctypes.POINTER(ctypes.c_uint32)]),
    "fa_filter_rad": (ctypes.c_int, [ctypes.POINTER(fa_box), ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.POINTER(fa_nlist), ctypes.c_int, ctypes.POINTER(fa_nlist),
                                     ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32)]),
    "fa_rdf_accumulate_query": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_float,
                                               ctypes.c_float, ctypes.c_int, ctypes.c_uint32, ctypes.c_float,
                                               ctypes.c_float, ctypes.c_void_p]),
    "fa_rdf_accumulate_self_noalloc": (ctypes.c_int, [ctypes.POINTER(fa_box), ctypes.c_void_p, ctypes.c_uint32,
                                                      ctypes.c_float, ctypes.c_float, ctypes.c_uint32,
                                                      ctypes.c_float, ctypes.c_float, ctypes.c_void_p]),
    "fa_local_density": (ctypes.c_int, [ctypes.POINTER(fa_nlist), ctypes.c_float, ctypes.c_float, ctypes.c_int,
                                        ctypes.c_void_p, ctypes.c_void_p]),
    "fa_pmftxy_accumulate": (ctypes.c_int, [ctypes.POINTER(fa_nlist), ctypes.c_void_p, ctypes.c_float,
                                            ctypes.c_float, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p]),
    "fa_bond_order_accumulate": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                                ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p]),
    "fa_bond_order_reduce": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                            ctypes.c_void_p]),
    "fa_correlation_accumulate": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                                 ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p]),
    "fa_correlation_reduce": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]),
    "fa_pmftxyt_accumulate": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float,
                                             ctypes.c_float, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                             ctypes.c_void_p]),
    "fa_pmftr12_accumulate": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float,
                                             ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p]),
    "fa_pmftxyz_accumulate": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                             ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_uint32,
                                             ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p]),
    "fa_pmft_reduce": (ctypes.c_int, [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                      ctypes.c_uint32, ctypes.c_void_p]),
    "fa_pmftxy_reduce": (ctypes.c_int, [ctypes.POINTER(fa_box), ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32,
                                        ctypes.c_float, ctypes.c_float, ctypes.c_uint32, ctypes.c_uint32,
                                        ctypes.c_uint32, ctypes.c_void_p]),
    "fa_rdf_bin_table": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_float, ctypes.c_float, ctypes.c_void_p,
                                        ctypes.c_uint64]),
    "fa_rdf_accumulate_nlist": (ctypes.c_int, [ctypes.POINTER(fa_nlist), ctypes.c_uint32, ctypes.c_float,
                                               ctypes.c_float, ctypes.c_void_p]),
    "fa_rdf_reduce": (ctypes.c_int, [ctypes.POINTER(fa_box), ctypes.c_void_p, ctypes.c_uint32, ctypes.c_float,
                                     ctypes.c_float, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                                     ctypes.c_void_p, ctypes.c_void_p]),
    "fa_steinhardt": (ctypes.c_int, [ctypes.POINTER(fa_box), ctypes.c_void_p, ctypes.c_uint32,
                                     ctypes.POINTER(fa_nlist), ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_void_p]),
    "fa_solidliquid": (ctypes.c_int, [ctypes.POINTER(fa_box), ctypes.c_void_p, ctypes.c_uint32,
                                      ctypes.POINTER(fa_nlist), ctypes.c_uint32, ctypes.c_float, ctypes.c_uint32,
                                      ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.POINTER(ctypes.c_uint32)]),
    "fa_cluster": (ctypes.c_int, [ctypes.c_uint32, ctypes.POINTER(fa_nlist), ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32)]),
    "fa_hexatic": (ctypes.c_int, [ctypes.POINTER(fa_box), ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(fa_nlist),
                                  ctypes.c_uint32, ctypes.c_int, ctypes.c_void_p]),
}


class LibraryMissing(RuntimeError):
    pass


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            msg = f"{LIB_PATH} is missing: build it with `python -m freud_amd.build`"
            raise LibraryMissing(msg)
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def exported_symbols():
    return sorted(_SIGS)


def check(rc):
    """Raise the exception type the reference raises for a C-ABI status."""
    if rc == FA_OK:
        return
    msg = lib().fa_last_error().decode(errors="replace")
    if rc == FA_ERR_INVALID:
        raise ValueError(msg)
    if rc == FA_ERR_RANGE:
        raise IndexError(msg)
    raise RuntimeError(msg)


def init():
    check(lib().fa_init())


def device_name():
    buf = ctypes.create_string_buffer(256)
    check(lib().fa_device_name(buf, 256))
    return buf.value.decode()


def profile_enable(on=True):
    check(lib().fa_profile_enable(int(bool(on))))


def profile_reset():
    check(lib().fa_profile_reset())


def profile_stats():
    """{kernel name: (launches, total_ms)} of the timed launches so far."""
    out = {}
    n = lib().fa_profile_count()
    buf = ctypes.create_string_buffer(256)
    for i in range(n):
        cnt = ctypes.c_uint64()
        ms = ctypes.c_double()
        check(lib().fa_profile_get(i, buf, 256, ctypes.byref(cnt), ctypes.byref(ms)))
        out[buf.value.decode()] = (int(cnt.value), float(ms.value))
    return out


def tune_enable(on=True):
    """Switch the autotuner on or off (on by default)."""
    check(lib().fa_tune_enable(int(bool(on))))


def tune_table():
    """The autotuner's table as {(analysis, key): variant}."""
    out = {}
    a, k, v = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    for i in range(lib().fa_tune_count()):
        check(lib().fa_tune_get(i, ctypes.byref(a), ctypes.byref(k), ctypes.byref(v)))
        out[(a.value, k.value)] = v.value
    return out


def tune_load(table):
    """Preset the autotuner's table (the dict tune_table returns)."""
    for (a, k), v in table.items():
        check(lib().fa_tune_set(int(a), int(k), int(v)))


def tune_reset():
    check(lib().fa_tune_reset())


def tune_variants(analysis, key):
    """Number of launch variants the autotuner chooses among for (analysis, key)."""
    n = lib().fa_tune_variants(int(analysis), int(key))
    check(min(n, 0))
    return n


def synchronize():
    check(lib().fa_synchronize())


class DeviceArray:
    """A device (HBM) allocation with a numpy dtype and shape."""

    __slots__ = ("ptr", "nbytes", "dtype", "shape", "__weakref__")

    def __init__(self, shape, dtype):
        init()
        self.dtype = np.dtype(dtype)
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        check(lib().fa_malloc(ctypes.byref(p), max(self.nbytes, 16)))
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, arr, dtype=None):
        a = np.ascontiguousarray(arr, dtype=dtype)
        d = cls(a.shape, a.dtype)
        if d.nbytes:
            check(lib().fa_memcpy_h2d(d.ptr, a.ctypes.data, d.nbytes))
        return d

    @classmethod
    def zeros(cls, shape, dtype):
        d = cls(shape, dtype)
        d.zero()
        return d

    def zero(self):
        if self.nbytes:
            check(lib().fa_memset(self.ptr, 0, self.nbytes))

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            check(lib().fa_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes))
        return out

    def __del__(self):
        p = getattr(self, "ptr", None)
        if p and _lib is not None:
            _lib.fa_free(p)
            self.ptr = None


def reuse(owner, name, shape, dtype):
    """owner.<name> as a DeviceArray of this shape and dtype, reusing the
    previous compute's allocation when it matches (results are copied to the
    host when read, so reuse never changes what an earlier read returned)."""
    shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
    a = getattr(owner, name, None)
    if a is None or a.shape != shape or a.dtype != np.dtype(dtype):
        a = DeviceArray(shape, dtype)
        setattr(owner, name, a)
    return a


def to_device_f32(arr):
    return DeviceArray.from_numpy(np.ascontiguousarray(arr, dtype=np.float32))


def copy_d2h(ptr, shape, dtype):
    out = np.empty(shape, dtype)
    if out.nbytes:
        check(lib().fa_memcpy_d2h(out.ctypes.data, ptr, out.nbytes))
    return out


def make_box(box):
    b = fa_box()
    b.Lx, b.Ly, b.Lz = float(box.Lx), float(box.Ly), float(box.Lz)
    b.xy, b.xz, b.yz = float(box.xy), float(box.xz), float(box.yz)
    b.is2D = int(bool(box.is2D))
    return b
