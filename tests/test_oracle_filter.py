"""Pins the oracle's NeighborList filters (orc_filter_sann / orc_filter_rad in
oracle/freud_oracle.cpp) against the reference's known answers
(tests/test_locality_filter.py: SANN_simple, RAD_simple, perfect crystals) and
against independent float64 statements of the two algorithms: the solid-angle
sum of van Meel et al. (2012) for SANN and the cosine form of the RAD blocking
test of Higham and Henchman (2016), which the reference's test file also uses."""

import numpy as np
import numpy.testing as npt
import pytest

from freud_amd import data
from freud_amd.box import Box


def _bt(box):
    return (box.Lx, box.Ly, box.Lz, box.xy, box.xz, box.yz, box.is2D)


def _ball(oracle, box, pts, r_max):
    return oracle.query(_bt(box), pts, pts, dict(mode="ball", r_max=r_max, exclude_ii=True))


def _wrap(box, v):
    # minimum image in a periodic box, float64
    b = Box.from_box(box)
    m = np.array(b.to_matrix(), np.float64)
    f = np.linalg.solve(m, np.asarray(v, np.float64).T).T
    f -= np.round(f)
    return f @ m.T


# --------------------------------------------------------------- known answers
def test_rad_simple_known_answer(oracle):
    # tests/test_locality_filter.py:200-257 of the reference
    pts = np.array([[0, 0, 0], [0, 2, 0], [0, 0, 4], [0, 1, 0], [0, 0, 2.1]], np.float32)
    box = Box.cube(10)
    nl = _ball(oracle, box, pts, 2.5)
    out, _ = oracle.filter_rad(_bt(box), pts, pts, nl, terminate_after_blocked=True)
    npt.assert_allclose(out["distances"], [1, 1, 1.9, 1, 1, 1.9, 2.1], rtol=1e-6)
    npt.assert_array_equal(out["ij"][:, 1], [3, 3, 4, 0, 1, 2, 0])
    npt.assert_array_equal(out["ij"][:, 0], [0, 1, 2, 3, 3, 4, 4])
    out, _ = oracle.filter_rad(_bt(box), pts, pts, nl, terminate_after_blocked=False)
    npt.assert_allclose(out["distances"], [1, 2.1, 1, 1.9, 1, 1, 1.9, 2.1], rtol=1e-6)
    npt.assert_array_equal(out["ij"][:, 1], [3, 4, 3, 4, 0, 1, 2, 0])
    npt.assert_array_equal(out["ij"][:, 0], [0, 0, 1, 2, 3, 3, 4, 4])


def test_sann_simple_known_answer(oracle):
    # tests/test_locality_filter.py:377-447 of the reference
    pts = np.array([[0, 0, 0], [1, 0, 0], [0, 2, 0], [0, 0, 4], [0, 1, 0], [0, 0, 1]], np.float32)
    box = Box.cube(10)
    nl = _ball(oracle, box, pts, 1.5)
    out, unfilled = oracle.filter_sann(nl, len(pts))
    s2 = np.sqrt(2)
    npt.assert_allclose(out["distances"], [1, 1, 1, 1, s2, s2, 1, 1, 1, s2, s2, 1, s2, s2], rtol=1e-6)
    npt.assert_array_equal(out["ij"][:, 1], [1, 4, 5, 0, 4, 5, 4, 0, 2, 1, 5, 0, 1, 4])
    npt.assert_array_equal(out["ij"][:, 0], [0, 0, 0, 1, 1, 1, 2, 4, 4, 4, 4, 5, 5, 5])
    # every point's list runs out before the solid angle fills (4 or fewer candidates)
    assert unfilled.all()


@pytest.mark.parametrize(("cls", "expected"), [("fcc", 12), ("bcc", 14)])
@pytest.mark.parametrize("which", ["sann", "rad_closed", "rad_open"])
def test_perfect_crystals(oracle, cls, expected, which):
    # tests/test_locality_filter.py:90-106 of the reference
    box, pts = getattr(data.UnitCell, cls)().generate_system(3)
    pts = np.asarray(pts, np.float32)
    nl = _ball(oracle, box, pts, 1.49)
    if which == "sann":
        out, unfilled = oracle.filter_sann(nl, len(pts))
    else:
        out, unfilled = oracle.filter_rad(_bt(box), pts, pts, nl, terminate_after_blocked=which == "rad_closed")
    counts = np.bincount(out["ij"][:, 0], minlength=len(pts))
    npt.assert_array_equal(counts, expected)
    assert not unfilled.any()


# ------------------------------------------------------ independent statements
def _sorted(nl):
    ij = nl["ij"]
    order = np.lexsort((nl["weights"], ij[:, 1], nl["distances"], ij[:, 0]))
    return ij[order], nl["distances"][order]


def _sann_float64(nl, n):
    """SANN neighbour count per point: the smallest m >= 3 with
    sum_{k<m} d_k / (m - 2) <= d_m (van Meel et al. 2012, eq. 3)."""
    ij, d = _sorted(nl)
    keep = []
    for i in range(n):
        seg = d[ij[:, 0] == i].astype(np.float64)
        m = 3
        while m < len(seg) and seg[:m].sum() / (m - 2) > seg[m]:
            m += 1
        keep.append(ij[ij[:, 0] == i][:m])
    return np.concatenate(keep)


def _rad_float64(box, pts, nl, closed, radii):
    """RAD: j is blocked by a closer k when
    r_j^2 / |v_j|^2 < r_k^2 cos(theta_jk) / |v_k|^2 (Higham and Henchman 2016)."""
    ij, _ = _sorted(nl)
    p = np.asarray(pts, np.float64)
    out = []
    for i in range(len(p)):
        nb = ij[ij[:, 0] == i, 1]
        v = _wrap(box, p[i] - p[nb])
        for jj in range(len(nb)):
            good = True
            for kk in range(jj):
                coz = v[jj] @ v[kk] / (np.linalg.norm(v[jj]) * np.linalg.norm(v[kk]))
                if radii[nb[jj]] ** 2 / (v[jj] @ v[jj]) < radii[nb[kk]] ** 2 * coz / (v[kk] @ v[kk]):
                    good = False
                    break
            if good:
                out.append((i, nb[jj]))
            elif closed:
                break
    return np.array(out, np.uint32).reshape(-1, 2)


def test_sann_random_system_matches_float64_statement(oracle):
    box, pts = data.make_random_system(10, 100, seed=1)
    pts = np.asarray(pts, np.float32)
    nl = _ball(oracle, box, pts, 4.9)
    out, unfilled = oracle.filter_sann(nl, len(pts))
    npt.assert_array_equal(out["ij"], _sann_float64(nl, len(pts)))
    assert not unfilled.any()
    # SANN keeps the unfiltered bonds' own vectors and weights
    ij, d = _sorted(nl)
    lookup = {(int(a), int(b)): k for k, (a, b) in enumerate(nl["ij"])}
    rows = [lookup[(int(a), int(b))] for a, b in out["ij"]]
    npt.assert_array_equal(out["vectors"], nl["vectors"][rows])
    npt.assert_array_equal(out["weights"], nl["weights"][rows])


@pytest.mark.parametrize("closed", [True, False])
@pytest.mark.parametrize("with_radii", [False, True])
def test_rad_random_system_matches_float64_statement(oracle, closed, with_radii):
    box, pts = data.make_random_system(10, 100, seed=1)
    pts = np.asarray(pts, np.float32)
    radii = np.linspace(0.75, 1.25, len(pts)).astype(np.float32) if with_radii else np.ones(len(pts), np.float32)
    nl = _ball(oracle, box, pts, 4.9)
    out, unfilled = oracle.filter_rad(_bt(box), pts, pts, nl, terminate_after_blocked=closed,
                                      points_radii=radii if with_radii else None)
    npt.assert_array_equal(out["ij"], _rad_float64(box, pts, nl, closed, radii))
    # RAD writes wrap(q_i - p_j) and weight 1 for the kept bonds
    v = _wrap(box, pts[out["ij"][:, 0]].astype(np.float64) - pts[out["ij"][:, 1]])
    npt.assert_allclose(out["vectors"], v, atol=2e-5)
    npt.assert_array_equal(out["weights"], 1.0)
    npt.assert_allclose(np.linalg.norm(out["vectors"], axis=1), out["distances"], rtol=2e-6)
    assert not unfilled.any()


def test_filtered_list_sorted_by_distance(oracle):
    box, pts = data.make_random_system(8, 60, seed=3)
    pts = np.asarray(pts, np.float32)
    nl = _ball(oracle, box, pts, 3.0)
    for out, _ in (oracle.filter_sann(nl, len(pts)), oracle.filter_rad(_bt(box), pts, pts, nl)):
        ij, d = out["ij"], out["distances"]
        key = np.lexsort((ij[:, 1], d, ij[:, 0]))
        npt.assert_array_equal(key, np.arange(len(d)))


def test_unfilled_shells(oracle):
    # tests/test_locality_filter.py:65-72 of the reference: too few candidates
    box, pts = data.make_random_system(10, 5, seed=1)
    pts = np.asarray(pts, np.float32)
    nl = _ball(oracle, box, pts, 1.2)
    counts = np.bincount(nl["ij"][:, 0], minlength=len(pts))
    _, unfilled = oracle.filter_sann(nl, len(pts))
    # fewer than four candidates can never close the SANN shell
    assert unfilled[counts < 4].all()
    assert unfilled.any()
    _, unfilled = oracle.filter_rad(_bt(box), pts, pts, nl, terminate_after_blocked=True)
    # RAD-closed: a point is unfilled when none of its candidates is blocked;
    # with at most one candidate nothing can be blocked
    assert unfilled[counts <= 1].all()
    _, unfilled_open = oracle.filter_rad(_bt(box), pts, pts, nl, terminate_after_blocked=False)
    assert not unfilled_open.any()


def test_sann_stops_where_the_sum_criterion_fails(oracle):
    # hand-built segment: d = 1, 1, 1, 1.4, 1.6, 5 -> m = 3: 3/1 > 1.4 keep,
    # m = 4: 4.4/2 = 2.2 > 1.6 keep, m = 5: 6.0/3 = 2 < 5 stop -> 5 kept
    d = np.array([1, 1, 1, 1.4, 1.6, 5], np.float32)
    nl = {"ij": np.array([[0, j + 1] for j in range(6)], np.uint32), "distances": d,
          "weights": np.full(6, 2.0, np.float32), "vectors": np.zeros((6, 3), np.float32)}
    out, unfilled = oracle.filter_sann(nl, 1)
    npt.assert_array_equal(out["ij"][:, 1], [1, 2, 3, 4, 5])
    npt.assert_array_equal(out["weights"], 2.0)
    assert not unfilled[0]
    assert out["num_query_points"] == 1 and out["num_points"] == 6


def test_sann_division_free_test_is_exact():
    # The device kernels test the SANN criterion RN(sum / (m - 2.0)) > d_m
    # (FilterSANN.cc:73, double division) as sum > d_m * (m - 2) in double,
    # exact for float sum, d_m and m - 2 < 2^29.  Check the two agree on
    # near-ties: sum within a few float spacings of d_m * k, and at random.
    rng = np.random.default_rng(0)
    n = 400_000
    d = rng.uniform(0.0, 8.0, n).astype(np.float32)
    d[:1000] = rng.uniform(0, 1e-30, 1000).astype(np.float32)  # tiny and subnormal-adjacent values
    k = rng.integers(1, 1 << 12, n)
    k[:5000] = (1 << 29) - 1 - rng.integers(0, 100, 5000)
    exact = (d.astype(np.float64) * k).astype(np.float32)
    step = rng.integers(-4, 5, n).astype(np.int32)
    s = (exact.view(np.int32) + step).view(np.float32)
    s[-50_000:] = rng.uniform(0.0, 3000.0, 50_000).astype(np.float32)
    ok = np.isfinite(s) & (s >= 0)
    s, d, k = s[ok], d[ok], k[ok]
    by_division = (s.astype(np.float64) / (k.astype(np.float64) + 2.0 - 2.0)) > d.astype(np.float64)
    by_product = s.astype(np.float64) > d.astype(np.float64) * k.astype(np.float64)
    npt.assert_array_equal(by_division, by_product)
    # the near-tie population holds both outcomes
    assert by_division.any() and (~by_division).any()
