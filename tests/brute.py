"""Numpy brute-force restatements of the reference's per-pair arithmetic, used
to check the oracle's tree and cell searches (test helpers).

All arithmetic is float32 with no fused multiply-add, in the operation order of
freud/locality/AABBQuery.cc:93,118-127 (AABB formula), LinkCell.cc:522-525
(LinkCell formula) and NeighborQuery.h:549-556 (image vectors).
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
