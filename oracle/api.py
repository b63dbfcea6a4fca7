"""ctypes wrapper of the CPU oracle (test infrastructure only).

Each function mirrors one freud entry point and returns numpy arrays in the
reference's layouts.  See ``freud_oracle.cpp`` for the citations of the
reference arithmetic each function restates.
"""

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "liboracle.so")
_lib = None

__all__ = [
    "build",
    "local_density",
    "pmftxy_accumulate",
    "pmftxy_reduce",
    "pmft3_accumulate",
    "correlation_accumulate",
    "bond_order",
    "pmft3_reduce",
    "box_op",
    "nearest_plane_distance",
    "query",
    "cellquery_grid",
    "all_pairs",
    "rdf_accumulate",
    "rdf_bin",
    "rdf_reduce",
    "regular_axis",
    "ylm",
    "wigner3j_table",
    "steinhardt",
    "solidliquid",
    "cluster",
    "hexatic",
    "filter_sann",
    "filter_rad",
    "OracleError",
]


class OracleError(Exception):
    pass


class _QueryParams(ctypes.Structure):
    _fields_ = [
        ("mode", ctypes.c_int),
        ("k", ctypes.c_uint32),
        ("r_max", ctypes.c_float),
        ("r_min", ctypes.c_float),
        ("r_guess", ctypes.c_float),
        ("scale", ctypes.c_float),
        ("exclude_ii", ctypes.c_int),
    ]


class _Nlist(ctypes.Structure):
    _fields_ = [
        ("ij", ctypes.POINTER(ctypes.c_uint32)),
        ("dist", ctypes.POINTER(ctypes.c_float)),
        ("w", ctypes.POINTER(ctypes.c_float)),
        ("vec", ctypes.POINTER(ctypes.c_float)),
        ("n", ctypes.c_uint64),
    ]


def build():
    """Compile the oracle library with its Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.orc_last_error.restype = ctypes.c_char_p
        _lib.orc_wigner3j_table.restype = ctypes.c_uint32
        _lib.orc_cluster.restype = ctypes.c_uint32
        _lib.orc_solidliquid.restype = ctypes.c_uint32
        _lib.orc_cellquery_grid.restype = ctypes.c_uint32
    return _lib


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None:
        a = a.reshape(shape)
    return a


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _boxp(box):
    """box: (Lx, Ly, Lz, xy, xz, yz, is2D) tuple or an object with those fields."""
    if hasattr(box, "Lx"):
        vals = [box.Lx, box.Ly, box.Lz, box.xy, box.xz, box.yz]
        is2d = bool(box.is2D)
    else:
        vals = list(box[:6])
        is2d = bool(box[6]) if len(box) > 6 else False
    return np.array(vals, dtype=np.float32), int(is2d)


def _check(rc):
    if rc == 0:
        return
    msg = _load().orc_last_error().decode()
    if rc == -1:
        raise ValueError(msg)
    raise RuntimeError(msg)


def box_op(box, vecs, op):
    """op: 'fractional' | 'absolute' | 'wrap' (Box.h:212-329)."""
    lib = _load()
    b, is2d = _boxp(box)
    v = _f32(vecs).reshape(-1, 3)
    out = np.empty_like(v)
    code = {"fractional": 0, "absolute": 1, "wrap": 2}[op]
    lib.orc_box_op(_p(b), is2d, code, _p(v), ctypes.c_uint64(len(v)), _p(out))
    return out


def nearest_plane_distance(box):
    lib = _load()
    b, is2d = _boxp(box)
    out = np.empty(3, np.float32)
    lib.orc_nearest_plane_distance(_p(b), is2d, _p(out))
    return out


def _qparams(mode=None, num_neighbors=None, r_max=None, r_min=None, r_guess=None, scale=None, exclude_ii=False):
    qp = _QueryParams()
    qp.mode = {None: 0, "ball": 1, "nearest": 2}[mode]
    qp.k = 0xFFFFFFFF if num_neighbors is None else int(num_neighbors)
    qp.r_max = -1.0 if r_max is None else float(r_max)
    qp.r_min = 0.0 if r_min is None else float(r_min)
    qp.r_guess = -1.0 if r_guess is None else float(r_guess)
    qp.scale = -1.0 if scale is None else float(scale)
    qp.exclude_ii = int(bool(exclude_ii))
    return qp


def _take_nlist(nl):
    lib = _load()
    n = int(nl.n)
    if n:
        ij = np.ctypeslib.as_array(nl.ij, shape=(n, 2)).copy()
        d = np.ctypeslib.as_array(nl.dist, shape=(n,)).copy()
        w = np.ctypeslib.as_array(nl.w, shape=(n,)).copy()
        v = np.ctypeslib.as_array(nl.vec, shape=(n, 3)).copy()
    else:
        ij = np.zeros((0, 2), np.uint32)
        d = np.zeros(0, np.float32)
        w = np.zeros(0, np.float32)
        v = np.zeros((0, 3), np.float32)
    lib.orc_nlist_free(ctypes.byref(nl))
    return {"ij": ij, "distances": d, "weights": w, "vectors": v}


def query(box, points, query_points, qargs, nq_type="aabb", cell_width=0.0, sort_by_distance=False, nthreads=1):
    """Ball / kNN query -> dict(ij, distances, weights, vectors) in NeighborList order.

    nq_type: 'aabb' (AABBQuery formula r = p_j - (q_i + image)), 'linkcell'
    (LinkCell formula r = box.wrap(p_j - q_i)) or 'cellquery' (CellQuery: ghosts
    near the faces, r = (p_j + shift) - q_i; ball queries only).
    """
    lib = _load()
    b, is2d = _boxp(box)
    pts = _f32(points).reshape(-1, 3)
    qpts = _f32(query_points).reshape(-1, 3)
    qp = _qparams(**qargs)
    nl = _Nlist()
    rc = lib.orc_query(
        _p(b), is2d, _p(pts), ctypes.c_uint32(len(pts)), _p(qpts), ctypes.c_uint32(len(qpts)),
        {"aabb": 0, "linkcell": 1, "cellquery": 2}[nq_type], ctypes.c_float(cell_width), qp, int(bool(sort_by_distance)),
        int(nthreads), ctypes.byref(nl))
    if rc:
        _check(rc)
    return _take_nlist(nl)


def cellquery_grid(box, points, r_cut):
    """CellQuery's grid for radius r_cut (CellQuery.h:136-160, CellQuery.cc:56-113):
    dict(n=(nx, ny, nz), min_pos, inverse_width, n_total, counts, counts_real),
    the per-cell arrays shaped (nz, ny, nx)."""
    lib = _load()
    b, is2d = _boxp(box)
    pts = _f32(points).reshape(-1, 3)
    n = np.zeros(3, np.int32)
    mn = np.zeros(3, np.float32)
    inv = ctypes.c_float()
    lib.orc_cellquery_grid(_p(b), is2d, _p(pts), ctypes.c_uint32(len(pts)), ctypes.c_float(r_cut), _p(n), _p(mn),
                           ctypes.byref(inv), None, None)
    ncell = int(n[0]) * int(n[1]) * int(n[2])
    counts = np.zeros(ncell, np.uint32)
    real = np.zeros(ncell, np.uint32)
    total = lib.orc_cellquery_grid(_p(b), is2d, _p(pts), ctypes.c_uint32(len(pts)), ctypes.c_float(r_cut), _p(n),
                                   _p(mn), ctypes.byref(inv), _p(counts), _p(real))
    shape = (int(n[2]), int(n[1]), int(n[0]))
    return dict(n=tuple(int(v) for v in n), min_pos=mn, inverse_width=np.float32(inv.value), n_total=int(total),
                counts=counts.reshape(shape), counts_real=real.reshape(shape))


def all_pairs(box, points, query_points, exclude_ii=True):
    lib = _load()
    b, is2d = _boxp(box)
    pts = _f32(points).reshape(-1, 3)
    qpts = _f32(query_points).reshape(-1, 3)
    nl = _Nlist()
    lib.orc_all_pairs(_p(b), is2d, _p(pts), ctypes.c_uint32(len(pts)), _p(qpts), ctypes.c_uint32(len(qpts)),
                      int(bool(exclude_ii)), ctypes.byref(nl))
    return _take_nlist(nl)


def regular_axis(bins, r_min, r_max):
    lib = _load()
    edges = np.empty(bins + 1, np.float32)
    centers = np.empty(bins, np.float32)
    w = ctypes.c_float()
    inv = ctypes.c_float()
    lib.orc_regular_axis(ctypes.c_uint32(bins), ctypes.c_float(r_min), ctypes.c_float(r_max), _p(edges),
                         _p(centers), ctypes.byref(w), ctypes.byref(inv))
    return edges, centers


def rdf_bin(distances, bins, r_min, r_max, counts=None):
    lib = _load()
    d = _f32(distances)
    if counts is None:
        counts = np.zeros(bins, np.uint32)
    lib.orc_rdf_bin(_p(d), ctypes.c_uint64(len(d)), ctypes.c_uint32(bins), ctypes.c_float(r_min),
                    ctypes.c_float(r_max), _p(counts))
    return counts


def rdf_accumulate(box, points, query_points, bins, r_max, r_min=0.0, exclude_ii=True, nq_type="aabb",
                   cell_width=0.0, query_r_max=None, query_r_min=0.0, counts=None, nthreads=1):
    """Fused ball query + RegularAxis binning (RDF.cc:101-110).  Adds to `counts`."""
    lib = _load()
    b, is2d = _boxp(box)
    pts = _f32(points).reshape(-1, 3)
    qpts = _f32(query_points).reshape(-1, 3)
    if counts is None:
        counts = np.zeros(bins, np.uint32)
    qrm = r_max if query_r_max is None else query_r_max
    rc = lib.orc_rdf_accumulate(
        _p(b), is2d, _p(pts), ctypes.c_uint32(len(pts)), _p(qpts), ctypes.c_uint32(len(qpts)),
        {"aabb": 0, "linkcell": 1, "cellquery": 2}[nq_type], ctypes.c_float(cell_width), ctypes.c_float(qrm),
        ctypes.c_float(query_r_min), int(bool(exclude_ii)), ctypes.c_uint32(bins), ctypes.c_float(r_min),
        ctypes.c_float(r_max), int(nthreads), _p(counts))
    _check(rc)
    return counts


def rdf_reduce(box, counts, bins, r_min, r_max, n_points, n_query_points, n_frames, finite_size=False):
    lib = _load()
    b, is2d = _boxp(box)
    c = np.ascontiguousarray(counts, dtype=np.uint32)
    rdf = np.empty(bins, np.float32)
    n_r = np.empty(bins, np.float32)
    lib.orc_rdf_reduce(_p(b), is2d, _p(c), ctypes.c_uint32(bins), ctypes.c_float(r_min), ctypes.c_float(r_max),
                       ctypes.c_uint32(n_points), ctypes.c_uint32(n_query_points), ctypes.c_uint32(n_frames),
                       int(bool(finite_size)), _p(rdf), _p(n_r))
    return rdf, n_r


def ylm(l, theta, phi, f64=True):
    """Y_l^m for each (theta, phi), m order [0..l, -1..-l]; complex128 (n, 2l+1)."""
    lib = _load()
    t = np.ascontiguousarray(theta, dtype=np.float64).ravel()
    p = np.ascontiguousarray(phi, dtype=np.float64).ravel()
    re = np.empty((len(t), 2 * l + 1))
    im = np.empty((len(t), 2 * l + 1))
    lib.orc_ylm(int(l), _p(t), _p(p), ctypes.c_uint64(len(t)), int(bool(f64)), _p(re), _p(im))
    return re + 1j * im


def wigner3j_table(l):
    lib = _load()
    n = lib.orc_wigner3j_table(int(l), None)
    out = np.empty(n, np.float64)
    lib.orc_wigner3j_table(int(l), _p(out))
    return out


def steinhardt(box, points, nlist, ls, average=False, wl=False, weighted=False, wl_normalize=False, f64=False,
               nthreads=1):
    """Steinhardt.cc:85-359 over an i-sorted NeighborList.

    Returns dict with per-l lists 'qlm' ((N, 2l+1) complex), 'qlm_ave', and
    arrays 'ql' (N, n_l), 'ql_ave', 'wl', 'order' (n_l,), 'particle_order'.
    """
    lib = _load()
    b, is2d = _boxp(box)
    pts = _f32(points).reshape(-1, 3)
    n = len(pts)
    ij = np.ascontiguousarray(nlist["ij"], dtype=np.uint32)
    d = _f32(nlist["distances"])
    w = _f32(nlist["weights"])
    ls = np.atleast_1d(np.asarray(ls, dtype=np.uint32))
    nl = len(ls)
    dt = np.float64 if f64 else np.float32
    total_m = int(np.sum(2 * ls + 1))
    qlm = np.zeros(2 * n * total_m, dt)
    qlm_ave = np.zeros(2 * n * total_m, dt)
    ql = np.zeros((n, nl), dt)
    ql_ave = np.zeros((n, nl), dt)
    wlv = np.zeros((n, nl), dt)
    order = np.zeros(nl, dt)
    lib.orc_steinhardt(_p(b), is2d, _p(pts), ctypes.c_uint32(n), _p(ij), _p(d), _p(w), ctypes.c_uint64(len(d)),
                       _p(ls), nl, int(average), int(wl), int(weighted), int(wl_normalize), int(f64), int(nthreads),
                       _p(qlm), _p(ql), _p(ql_ave), _p(wlv), _p(order), _p(qlm_ave))
    out_qlm, out_ave = [], []
    off = 0
    for l in ls:
        nm = 2 * int(l) + 1
        blk = qlm[2 * n * off: 2 * n * (off + nm)].reshape(n, nm, 2)
        out_qlm.append(blk[..., 0] + 1j * blk[..., 1])
        blk = qlm_ave[2 * n * off: 2 * n * (off + nm)].reshape(n, nm, 2)
        out_ave.append(blk[..., 0] + 1j * blk[..., 1])
        off += nm
    if wl:
        po = wlv
    elif average:
        po = ql_ave
    else:
        po = ql
    return {"qlm": out_qlm, "qlm_ave": out_ave, "ql": ql, "ql_ave": ql_ave, "wl": wlv, "order": order,
            "particle_order": po}


def solidliquid(nlist, qlm, ql, l, q_threshold, solid_threshold, normalize_q=True):
    """SolidLiquid.cc:31-114 given the Steinhardt(l) q_lm / q_l of the same nlist."""
    lib = _load()
    ij = np.ascontiguousarray(nlist["ij"], dtype=np.uint32)
    nb = len(ij)
    q = np.ascontiguousarray(qlm, dtype=np.complex64)
    n = q.shape[0]
    qlv = _f32(ql).ravel()
    ql_ij = np.empty(nb, np.float32)
    nc = np.empty(n, np.uint32)
    cidx = np.empty(n, np.uint32)
    ncl = lib.orc_solidliquid(ctypes.c_uint32(n), _p(ij), ctypes.c_uint64(nb), int(l), _p(q), _p(qlv),
                              ctypes.c_float(q_threshold), ctypes.c_uint32(int(solid_threshold)),
                              int(bool(normalize_q)), _p(ql_ij), _p(nc), _p(cidx))
    return {"ql_ij": ql_ij, "num_connections": nc, "cluster_idx": cidx, "num_clusters": int(ncl)}


def cluster(n_points, ij):
    lib = _load()
    ij = np.ascontiguousarray(ij, dtype=np.uint32).reshape(-1, 2)
    cidx = np.empty(n_points, np.uint32)
    ncl = lib.orc_cluster(ctypes.c_uint32(n_points), _p(ij), ctypes.c_uint64(len(ij)), _p(cidx))
    return cidx, int(ncl)


def hexatic(box, points, nlist, k=6, weighted=False):
    lib = _load()
    b, is2d = _boxp(box)
    pts = _f32(points).reshape(-1, 3)
    ij = np.ascontiguousarray(nlist["ij"], dtype=np.uint32)
    w = _f32(nlist["weights"])
    psi = np.empty(2 * len(pts), np.float32)
    lib.orc_hexatic(_p(b), is2d, _p(pts), ctypes.c_uint32(len(pts)), _p(ij), _p(w), ctypes.c_uint64(len(ij)),
                    ctypes.c_uint32(k), int(bool(weighted)), _p(psi))
    return psi.reshape(-1, 2)[:, 0] + 1j * psi.reshape(-1, 2)[:, 1]


def local_density(nlist, n_query_points, r_max, diameter, is2D=False):
    """LocalDensity.cc:39-84 over an i-sorted NeighborList -> (density, num_neighbors)."""
    lib = _load()
    ij = np.ascontiguousarray(nlist["ij"], dtype=np.uint32)
    d = _f32(nlist["distances"])
    dens = np.empty(n_query_points, np.float32)
    num = np.empty(n_query_points, np.float32)
    lib.orc_local_density(ctypes.c_uint32(n_query_points), _p(ij), _p(d), ctypes.c_uint64(len(d)),
                          ctypes.c_float(r_max), ctypes.c_float(diameter), int(bool(is2D)), _p(dens), _p(num))
    return dens, num


def pmftxy_accumulate(nlist, query_angles, x_max, y_max, n_x, n_y, counts=None):
    """PMFTXY.cc:69-88: adds the rotated-bond histogram (n_x, n_y) to counts."""
    lib = _load()
    ij = np.ascontiguousarray(nlist["ij"], dtype=np.uint32)
    v = _f32(nlist["vectors"]).reshape(-1, 3)
    th = _f32(query_angles).ravel()
    if counts is None:
        counts = np.zeros((n_x, n_y), np.uint32)
    lib.orc_pmftxy_accumulate(_p(ij), _p(v), ctypes.c_uint64(len(ij)), _p(th), ctypes.c_float(x_max),
                              ctypes.c_float(y_max), ctypes.c_uint32(n_x), ctypes.c_uint32(n_y), _p(counts))
    return counts


def pmftxy_reduce(box, counts, x_max, y_max, n_points, n_query_points, n_frames):
    lib = _load()
    b, is2d = _boxp(box)
    c = np.ascontiguousarray(counts, dtype=np.uint32)
    n_x, n_y = c.shape
    pcf = np.empty((n_x, n_y), np.float32)
    lib.orc_pmftxy_reduce(_p(b), is2d, _p(c), ctypes.c_uint32(n_x), ctypes.c_uint32(n_y), ctypes.c_float(x_max),
                          ctypes.c_float(y_max), ctypes.c_uint32(n_points), ctypes.c_uint32(n_query_points),
                          ctypes.c_uint32(n_frames), _p(pcf))
    return pcf


_PMFT3_KIND = {"xyt": 0, "r12": 1, "xyz": 2}


def _pmft3_axes(kind, maxes, bins):
    two_pi = np.float32(2.0 * np.pi)
    m = [np.float32(v) for v in maxes]
    if kind == "xyt":
        return [(-m[0], m[0]), (-m[1], m[1]), (np.float32(0), two_pi)]
    if kind == "r12":
        return [(np.float32(0), m[0]), (np.float32(0), two_pi), (np.float32(0), two_pi)]
    return [(-m[0], m[0]), (-m[1], m[1]), (-m[2], m[2])]


def pmft3_accumulate(kind, nl, maxes, bins, p_theta=None, q_theta=None, q_quat=None, equiv=None, counts=None):
    """PMFTXYT ('xyt'), PMFTR12 ('r12') or PMFTXYZ ('xyz') bin counts of a
    NeighborList dict (PMFTXYT.cc:76-100, PMFTR12.cc:90-111, PMFTXYZ.cc:106-145)."""
    lib = _load()
    ax = _pmft3_axes(kind, maxes, bins)
    n = np.array(bins, np.uint32)
    lo = np.array([a[0] for a in ax], np.float32)
    hi = np.array([a[1] for a in ax], np.float32)
    if counts is None:
        counts = np.zeros(int(np.prod(n)), np.uint32)
    ij = np.ascontiguousarray(nl["ij"], np.uint32)
    vec = _f32(nl["vectors"])
    dist = _f32(nl["distances"])
    pt = _f32(p_theta) if p_theta is not None else np.zeros(1, np.float32)
    qt = _f32(q_theta) if q_theta is not None else np.zeros(1, np.float32)
    qq = _f32(q_quat) if q_quat is not None else np.zeros(4, np.float32)
    eq = _f32(equiv) if equiv is not None else np.array([1, 0, 0, 0], np.float32)
    n_equiv = eq.size // 4
    lib.orc_pmft3_accumulate(_PMFT3_KIND[kind], _p(ij), _p(vec), _p(dist), ctypes.c_uint64(len(dist)), _p(pt),
                             _p(qt), _p(qq), _p(eq), ctypes.c_uint32(n_equiv), _p(n), _p(lo), _p(hi), _p(counts))
    return counts.reshape(tuple(int(v) for v in n))


def pmft3_reduce(kind, box, counts, maxes, n_points, n_query_points, n_frames, n_equiv=1):
    lib = _load()
    b, is2d = _boxp(box)
    c = np.ascontiguousarray(counts, dtype=np.uint32)
    n = np.array(c.shape, np.uint32)
    mx = np.array(list(maxes) + [0.0] * (3 - len(maxes)), np.float32)
    pcf = np.empty(c.size, np.float32)
    lib.orc_pmft3_reduce(_PMFT3_KIND[kind], _p(b), is2d, _p(c), _p(n), _p(mx), ctypes.c_uint32(n_points),
                         ctypes.c_uint32(n_query_points), ctypes.c_uint32(n_frames), ctypes.c_uint32(n_equiv),
                         _p(pcf))
    return pcf.reshape(c.shape)


def correlation_accumulate(nl, values, query_values, bins, r_max):
    """CorrelationFunction (CorrelationFunction.cc:50-98) over a NeighborList
    dict: (counts, correlation) with correlation = sums / counts."""
    lib = _load()
    ij = np.ascontiguousarray(nl["ij"], np.uint32)
    dist = _f32(nl["distances"])
    v = np.ascontiguousarray(values, np.complex128)
    q = np.ascontiguousarray(query_values, np.complex128)
    counts = np.zeros(bins, np.uint32)
    sums = np.zeros(2 * bins, np.float64)
    lib.orc_correlation_accumulate(_p(ij), _p(dist), ctypes.c_uint64(len(dist)), _p(v), _p(q), ctypes.c_uint32(bins),
                                   ctypes.c_float(r_max), _p(counts), _p(sums))
    # complex<double> /= count divides each part by the count (a scalar)
    nz = counts != 0
    s2 = sums.reshape(-1, 2).copy()
    s2[nz] = s2[nz] / counts[nz, None].astype(np.float64)
    return counts, s2.reshape(-1).view(np.complex128)


_BOND_ORDER_MODES = {"bod": 0, "lbod": 1, "obcd": 2, "oocd": 3}


def bond_order(nl, p_quat, q_quat, mode, n_theta, n_phi, n_frames=1, counts=None):
    """BondOrder (BondOrder.cc:26-136) over a NeighborList dict: (counts, bond_order)."""
    lib = _load()
    ij = np.ascontiguousarray(nl["ij"], np.uint32)
    vec = _f32(nl["vectors"])
    pq = _f32(p_quat)
    qq = _f32(q_quat)
    if counts is None:
        counts = np.zeros(n_theta * n_phi, np.uint32)
    lib.orc_bond_order_accumulate(_p(ij), _p(vec), ctypes.c_uint64(len(ij)), _p(pq), _p(qq),
                                  _BOND_ORDER_MODES[mode], ctypes.c_uint32(n_theta), ctypes.c_uint32(n_phi),
                                  _p(counts))
    bo = np.empty(n_theta * n_phi, np.float32)
    lib.orc_bond_order_reduce(_p(counts), ctypes.c_uint32(n_theta), ctypes.c_uint32(n_phi),
                              ctypes.c_uint32(n_frames), _p(bo))
    return counts.reshape(n_theta, n_phi), bo.reshape(n_theta, n_phi)


def _sorted_by_distance(nlist):
    """NeighborList::sort(true) (NeighborList.cc:371-399): order (i, d, j, w)."""
    ij = np.ascontiguousarray(nlist["ij"], dtype=np.uint32).reshape(-1, 2)
    d = _f32(nlist["distances"])
    w = _f32(nlist["weights"])
    v = _f32(nlist["vectors"]).reshape(-1, 3)
    order = np.lexsort((w, ij[:, 1], d, ij[:, 0]))
    return (np.ascontiguousarray(ij[order]), np.ascontiguousarray(d[order]), np.ascontiguousarray(w[order]),
            np.ascontiguousarray(v[order]))


def _filtered(ij, d, w, v, keep, unfilled):
    """The NeighborList the filters build from their bonds (NeighborList.cc:130-173):
    num_query_points / num_points are one past the largest kept indices."""
    keep = keep.astype(bool)
    out = {"ij": ij[keep], "distances": d[keep], "weights": w[keep], "vectors": v[keep]}
    if keep.any():
        out["num_query_points"] = int(out["ij"][:, 0].max()) + 1
        out["num_points"] = int(out["ij"][:, 1].max()) + 1
    return out, unfilled.astype(bool)


def filter_sann(nlist, num_query_points):
    """FilterSANN.cc:22-104 -> (filtered nlist dict, unfilled mask per query point)."""
    lib = _load()
    ij, d, w, v = _sorted_by_distance(nlist)
    keep = np.zeros(len(d), np.uint8)
    unfilled = np.zeros(num_query_points, np.uint8)
    lib.orc_filter_sann(ctypes.c_uint32(num_query_points), _p(ij), _p(d), ctypes.c_uint64(len(d)), _p(keep),
                        _p(unfilled))
    return _filtered(ij, d, w, v, keep, unfilled)


def filter_rad(box, points, query_points, nlist, terminate_after_blocked=True, points_radii=None):
    """FilterRAD.cc:22-131 -> (filtered nlist dict with vectors wrap(q_i - p_j) and
    weights 1, unfilled mask per query point)."""
    lib = _load()
    b, is2d = _boxp(box)
    pts = _f32(points).reshape(-1, 3)
    qpts = _f32(query_points).reshape(-1, 3)
    radii = None if points_radii is None else _f32(points_radii).reshape(-1)
    ij, d, w, v = _sorted_by_distance(nlist)
    keep = np.zeros(len(d), np.uint8)
    vec = np.zeros((len(d), 3), np.float32)
    unfilled = np.zeros(len(qpts), np.uint8)
    lib.orc_filter_rad(_p(b), is2d, _p(pts), _p(qpts), None if radii is None else _p(radii),
                       ctypes.c_uint32(len(qpts)), _p(ij), _p(d), ctypes.c_uint64(len(d)),
                       int(bool(terminate_after_blocked)), _p(keep), _p(vec), _p(unfilled))
    return _filtered(ij, d, np.ones_like(w), vec, keep, unfilled)
