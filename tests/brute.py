"""Numpy brute-force restatements of the reference's per-pair arithmetic, used
to check the oracle's tree and cell searches (test helpers).

All arithmetic is float32 with no fused multiply-add, in the operation order of
freud/locality/AABBQuery.cc:93,118-127 (AABB formula), LinkCell.cc:522-525
(LinkCell formula), NeighborQuery.h:549-556 (image vectors) and
CellQuery.h:136-298 / CellIterator.h:94-187 (CellQuery ghosts and grid).
"""

import numpy as np

f32 = np.float32


def image_vectors(box):
    """27 (9 in 2D) image vectors float(i)*a + float(j)*b + float(k)*c."""
    L = np.array([box.Lx, box.Ly, 0 if box.is2D else box.Lz], dtype=f32)
    a = np.array([L[0], 0, 0], dtype=f32)
    b = np.array([L[1] * f32(box.xy), L[1], 0], dtype=f32)
    c = np.zeros(3, f32) if box.is2D else np.array([L[2] * f32(box.xz), L[2] * f32(box.yz), L[2]], dtype=f32)
    out = []
    for i in (-1, 0, 1):
        for j in (-1, 0, 1):
            for k in (-1, 0, 1):
                if box.is2D and k != 0:
                    continue
                v = (f32(i) * a + f32(j) * b).astype(f32) + f32(k) * c
                out.append(v.astype(f32))
    return np.array(out, dtype=f32)


def dot3(r):
    return ((r[..., 0] * r[..., 0] + r[..., 1] * r[..., 1]) + r[..., 2] * r[..., 2]).astype(f32)


def ball_aabb(box, points, query_points, r_max, r_min=0.0, exclude_ii=False):
    """All (i, j, image) with r = p_j - (q_i + image), r_sq < r_max^2, r_sq >= r_min^2."""
    p = np.asarray(points, f32).copy()
    q = np.asarray(query_points, f32).copy()
    if box.is2D:
        p[:, 2] = 0
        q[:, 2] = 0
    rmax2 = f32(r_max) * f32(r_max)
    rmin2 = f32(r_min) * f32(r_min)
    bonds = []
    for img in image_vectors(box):
        qi = (q + img).astype(f32)
        r = (p[None, :, :] - qi[:, None, :]).astype(f32)
        rsq = dot3(r)
        ii, jj = np.nonzero((rsq < rmax2) & (rsq >= rmin2))
        for i, j in zip(ii, jj):
            if exclude_ii and i == j:
                continue
            bonds.append((i, j, np.sqrt(rsq[i, j]), r[i, j]))
    return _to_nlist(bonds)


def ball_linkcell(box, points, query_points, r_max, r_min=0.0, exclude_ii=False):
    """All (i, j) with r = box.wrap(p_j - q_i), r_sq < r_max^2, r_sq >= r_min^2."""
    p = np.asarray(points, f32)
    q = np.asarray(query_points, f32)
    rmax2 = f32(r_max) * f32(r_max)
    rmin2 = f32(r_min) * f32(r_min)
    d = (p[None, :, :] - q[:, None, :]).astype(f32).reshape(-1, 3)
    r = box.wrap(d).reshape(len(q), len(p), 3)
    rsq = dot3(r)
    bonds = []
    ii, jj = np.nonzero((rsq < rmax2) & (rsq >= rmin2))
    for i, j in zip(ii, jj):
        if exclude_ii and i == j:
            continue
        bonds.append((i, j, np.sqrt(rsq[i, j]), r[i, j]))
    return _to_nlist(bonds)


def cellquery_min_pos(box, r_cut):
    """Lower grid corner (box_min_x - r, box_min_y - r, -r) + makeAbsolute(0, 0, 0)."""
    r = f32(r_cut)
    Ly = f32(box.Ly)
    Lz = f32(0 if box.is2D else box.Lz)
    bmx = f32(Ly * min(f32(0), f32(box.xy))) + f32(Lz * min(f32(0), f32(box.xz)))
    bmy = f32(Lz * min(f32(0), f32(box.yz)))
    a0 = box.make_absolute(np.zeros(3, f32))
    return np.array([f32(bmx - r) + a0[0], f32(bmy - r) + a0[1], f32(-r) + a0[2]], dtype=f32)


def cellquery_entries(box, points, r_cut):
    """Every entry of CellQuery's grid as (point index, position): the points
    and each ghost p + shift whose shift generateGhosts makes (components s_a
    in {-1, 0, 1}, each nonzero one equal to the point's face flag
    d_a = [f_a <= r/npd_a] - [f_a >= 1 - r/npd_a], d_z = 0 in 2D), all kept
    only when their cell lies in the grid.  Brute force over all 27 shifts."""
    p = np.asarray(points, f32)
    r = f32(r_cut)
    npd = box.nearest_plane_distance()
    f = box.make_fractional(p)
    frc = np.array([r / f32(npd[0]), r / f32(npd[1]), (r / f32(npd[2])) if not box.is2D else f32(np.inf)],
                   dtype=f32)
    d = (f <= frc).astype(int) - (f.astype(np.float64) >= 1.0 - frc.astype(np.float64)).astype(int)
    if box.is2D:
        d[:, 2] = 0
    Lx, Ly = f32(box.Lx), f32(box.Ly)
    Lz = f32(0 if box.is2D else box.Lz)
    lat = [np.array([Lx, 0, 0], f32), np.array([Ly * f32(box.xy), Ly, 0], f32),
           np.zeros(3, f32) if box.is2D else np.array([Lz * f32(box.xz), Lz * f32(box.yz), Lz], f32)]
    idx, pos = [np.arange(len(p))], [p]
    for s in np.ndindex(3, 3, 3):
        s = np.array(s) - 1
        if not s.any():
            continue
        ok = np.all((s == 0) | (s[None, :] == d), axis=1)
        sh = np.zeros(3, f32)
        for a in range(3):
            if s[a]:
                sh = (sh + (lat[a] if s[a] == 1 else -lat[a])).astype(f32)
        idx.append(np.nonzero(ok)[0])
        pos.append((p[ok] + sh).astype(f32))
    idx = np.concatenate(idx)
    pos = np.concatenate(pos).astype(f32)
    cell = cellquery_cells(box, pos, r)
    inside = np.all(cell >= 0, axis=1) & np.all(cell < np.array(cellquery_dims(box, r)), axis=1)
    return idx[inside], pos[inside]


def cellquery_dims(box, r_cut):
    inv = f32(1) / f32(r_cut)
    Lx, Ly = f32(box.Lx), f32(box.Ly)
    Lz = f32(0 if box.is2D else box.Lz)
    wx = f32(f32(Lx + f32(Ly * abs(f32(box.xy)))) + f32(Lz * abs(f32(box.xz))))
    wy = f32(Ly + f32(Lz * abs(f32(box.yz))))
    return (int(f32(wx * inv)) + 3, int(f32(wy * inv)) + 3, int(f32(Lz * inv)) + 3)


def cellquery_cells(box, pos, r_cut):
    inv = f32(1) / f32(r_cut)
    mn = cellquery_min_pos(box, r_cut)
    return np.floor(((np.asarray(pos, f32) - mn).astype(f32) * inv).astype(f32)).astype(np.int64)


def ball_cellquery(box, points, query_points, r_max, r_min=0.0, exclude_ii=False):
    """All (i, entry) with the entry's cell within one cell (Chebyshev) of query
    point i's cell, i's cell inside the grid, r = entry - q_i,
    r_sq < r_max^2 and r_sq >= r_min^2."""
    j, pos = cellquery_entries(box, points, r_max)
    q = np.asarray(query_points, f32)
    dims = np.array(cellquery_dims(box, r_max))
    cq = cellquery_cells(box, q, r_max)
    ce = cellquery_cells(box, pos, r_max)
    q_in = np.all(cq >= 0, axis=1) & np.all(cq < dims, axis=1)
    near = np.all(np.abs(cq[:, None, :] - ce[None, :, :]) <= 1, axis=2) & q_in[:, None]
    r = (pos[None, :, :] - q[:, None, :]).astype(f32)
    rsq = dot3(r)
    rmax2 = f32(r_max) * f32(r_max)
    rmin2 = f32(r_min) * f32(r_min)
    keep = near & (rsq < rmax2) & (rsq >= rmin2)
    if exclude_ii:
        keep &= j[None, :] != np.arange(len(q))[:, None]
    bonds = []
    for i, e in zip(*np.nonzero(keep)):
        bonds.append((i, j[e], np.sqrt(rsq[i, e]), r[i, e]))
    return _to_nlist(bonds)


def knn_aabb(box, points, query_points, k, exclude_ii=False, r_max=np.inf, r_min=0.0):
    """k nearest by the AABB formula, closest image per point, ties by index.
    A point whose closest image lies nearer than r_min (d < r_min) is left
    out (AABBQuery.cc:188-205, 235-244)."""
    full = ball_aabb(box, points, query_points, r_max=min(r_max, 1e30), exclude_ii=exclude_ii)
    return _select_k(full, k, r_min)


def knn_linkcell(box, points, query_points, k, exclude_ii=False, r_max=np.inf, r_min=0.0):
    """k nearest by the LinkCell formula; r_min as r_sq >= r_min^2 (LinkCell.cc:616-630)."""
    full = ball_linkcell(box, points, query_points, r_max=min(r_max, 1e30), r_min=r_min, exclude_ii=exclude_ii)
    return _select_k(full, k)


def _select_k(full, k, r_min=0.0):
    ij, d, v = full["ij"], full["distances"], full["vectors"]
    best = {}
    for t in range(len(ij)):
        key = (int(ij[t, 0]), int(ij[t, 1]))
        if key not in best or d[t] < d[best[key]]:
            best[key] = t
    per_i = {}
    for (i, j), t in best.items():
        if not d[t] >= f32(r_min):
            continue
        per_i.setdefault(i, []).append((d[t], j, t))
    keep = []
    for i, lst in per_i.items():
        lst.sort()
        keep += [t for _, _, t in lst[:k]]
    bonds = [(ij[t, 0], ij[t, 1], d[t], v[t]) for t in keep]
    return _to_nlist(bonds)


def _to_nlist(bonds):
    bonds.sort(key=lambda b: (int(b[0]), int(b[1]), float(b[2])))
    if not bonds:
        return {"ij": np.zeros((0, 2), np.uint32), "distances": np.zeros(0, f32), "vectors": np.zeros((0, 3), f32)}
    return {
        "ij": np.array([(b[0], b[1]) for b in bonds], dtype=np.uint32),
        "distances": np.array([b[2] for b in bonds], dtype=f32),
        "vectors": np.array([b[3] for b in bonds], dtype=f32),
    }


def distance_tie_points(box, formula="aabb", radius=1.3, seed=0, tries=200000):
    """Two points A, B around a query point at the origin whose bond r_sq
    differ by one float step while sqrtf gives both the same distance d: a
    selection by (r_sq, j) and one by (d, j) then keep different points when
    B, the farther in r_sq, has the lower index.  Returns (A, B, r_sq_A,
    r_sq_B, d) with A, B float32 (3,) positions; the bond vector of each is
    p - 0 (AABB formula, image 0) or box.wrap(p) (LinkCell formula)."""
    rng = np.random.default_rng(seed)

    def rsq_of(p):
        v = p[None, :] if formula == "aabb" else box.wrap(p[None, :])
        return dot3(v.astype(f32))[0]

    ang = rng.random(tries) * 2 * np.pi
    cand = np.zeros((tries, 3), f32)
    cand[:, 0] = (radius * np.cos(ang)).astype(f32)
    cand[:, 1] = (radius * np.sin(ang)).astype(f32)
    r = np.array([rsq_of(c) for c in cand[:4000]], f32)
    by_rsq = {}
    for c, x in zip(cand[:4000], r):
        by_rsq.setdefault(x.view(np.uint32).item(), c)
    for bits, a in sorted(by_rsq.items()):
        nb = bits + 1
        if nb in by_rsq:
            ra, rb = np.uint32(bits).view(f32), np.uint32(nb).view(f32)
            if np.sqrt(ra) == np.sqrt(rb):
                return a, by_rsq[nb], ra, rb, np.sqrt(ra)
    raise RuntimeError("no distance tie found")


def knn_tie_system(box, formula="aabb", n_fill=300, seed=0):
    """Points for a kNN distance tie at the k-th slot of query point 0 (at the
    origin): 1 = B and 2 = A of distance_tie_points (equal d, r_sq of B one
    step above A's), 3 and 4 nearer (d = 0.5 and 0.8), then n_fill points
    farther than 4 from the origin.  Query 0 keeps {B, 3, 4} for k = 3 under
    the (d, j) rule; an (r_sq, j) rule would keep A instead of B."""
    a, b, ra, rb, d = distance_tie_points(box, formula, seed=seed)
    rng = np.random.default_rng(seed + 1)
    fill = []
    while len(fill) < n_fill:
        p = ((rng.random(3) - 0.5) * np.array([box.Lx, box.Ly, box.Lz], f32)).astype(f32)
        if np.linalg.norm(p) > 4.0:
            fill.append(p)
    pts = np.concatenate([np.zeros((1, 3), f32), b[None], a[None], np.array([[0, 0, 0.5], [0, 0, -0.8]], f32),
                          np.array(fill, f32)]).astype(f32)
    return pts
