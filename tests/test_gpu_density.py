"""GPU parity of LocalDensity and PMFTXY (the other consumers of the query
path): the HIP kernels through the C-ABI against the oracle on the same
NeighborLists, plus the reference's known answers (tests/test_density_local_density.py,
tests/test_pmft.py of freud 3.5.0)."""

import numpy as np
import numpy.testing as npt
import pytest

import freud_amd
from freud_amd.box import Box
from freud_amd.pmft import PMFTXY

pytestmark = pytest.mark.gpu


def _bt(box):
    return (box.Lx, box.Ly, box.Lz, box.xy, box.xz, box.yz, box.is2D)


def test_local_density_known_answer():
    # test_density_local_density.py:80-98 of the reference
    box = Box.cube(10)
    points = np.array([[0, 0, 0], [1, 0, 0]], np.float32)
    query_points = np.array([[0, 1, 0], [-1, -1, 0], [4, 0, 0]], np.float32)
    ld = freud_amd.density.LocalDensity(2, 1)
    with pytest.raises(AttributeError):
        ld.density
    ld.compute((box, points), query_points)
    v = 4 / 3 * 8 * np.pi
    d1 = np.linalg.norm(points[1] - query_points[1])
    frac = 1 if d1 < 1.5 else (0 if d1 > 2.5 else -d1 + 2 + 0.5)
    npt.assert_allclose(ld.density, [2 / v, (1 + frac) / v, 0], rtol=1e-4)
    assert ld.box == box
    with pytest.raises(ValueError):
        freud_amd.density.LocalDensity(-1, 10)
    with pytest.raises(ValueError):
        freud_amd.density.LocalDensity(2, -1)


@pytest.mark.parametrize("is2D", [False, True])
def test_local_density_matches_oracle(oracle, is2D):
    box, pos = freud_amd.data.make_random_system(10, 10000, is2D=is2D, seed=123)
    pos = pos.astype(np.float32)
    r_max, diameter = (3.0, 1.0) if not is2D else (2.0, 0.5)
    ld = freud_amd.density.LocalDensity(r_max, diameter).compute((box, pos))
    ref_nl = oracle.query(_bt(box), pos, pos, dict(r_max=r_max + 0.5 * diameter, exclude_ii=True), nthreads=8)
    dens, num = oracle.local_density(ref_nl, len(pos), r_max, diameter, is2D)
    npt.assert_array_equal(ld.num_neighbors, num)
    npt.assert_array_equal(ld.density, dens)
    if not is2D:
        npt.assert_array_less(np.abs(ld.density - 10.0), 1.5)
    # a given NeighborList gives the same sums
    nl = freud_amd.AABBQuery(box, pos).query(pos, dict(r_max=r_max + 0.5 * diameter, exclude_ii=True)).toNeighborList()
    ld2 = freud_amd.density.LocalDensity(r_max, diameter).compute((box, pos), neighbors=nl)
    npt.assert_array_equal(ld2.density, dens)


def test_pmftxy_matches_oracle(oracle):
    box, pts = freud_amd.data.make_random_system(10, 2500, is2D=True, seed=3)
    pts = pts.astype(np.float32)
    rng = np.random.default_rng(5)
    angles = (rng.random(len(pts)) * 2 * np.pi).astype(np.float32)
    lim = np.float32(np.sqrt(9 / 2))
    pm = PMFTXY(lim, lim, (100, 90))
    pm.compute((box, pts), angles)
    ref_nl = oracle.query(_bt(box), pts, pts, dict(r_max=float(pm.r_max), exclude_ii=True), nthreads=8)
    counts = oracle.pmftxy_accumulate(ref_nl, angles, lim, lim, 100, 90)
    npt.assert_array_equal(pm.bin_counts, counts)
    pcf = oracle.pmftxy_reduce(_bt(box), counts, lim, lim, len(pts), len(pts), 1)
    npt.assert_array_equal(pm._pcf, pcf)
    with np.errstate(divide="ignore"):
        npt.assert_array_equal(pm.pmft, -np.log(pcf))
    # reset=False accumulates a second frame; the normalisation counts frames
    pm.compute((box, pts), angles, reset=False)
    npt.assert_array_equal(pm.bin_counts, 2 * counts)
    npt.assert_array_equal(pm._pcf, oracle.pmftxy_reduce(_bt(box), 2 * counts, lim, lim, len(pts), len(pts), 2))
    # bin edges and centres (test_pmft.py:385-392)
    npt.assert_allclose(pm.bin_edges[0], np.linspace(-lim, lim, 101, dtype=np.float32), atol=1e-5)
    npt.assert_allclose(pm.bin_edges[1], np.linspace(-lim, lim, 91, dtype=np.float32), atol=1e-5)
    assert pm.nbins == [100, 90]


def test_pmftxy_known_answers():
    # test_pmft.py:455-485: every query point finds the origin as its nearest neighbour
    box = Box.square(8)
    points = np.array([[0, 0, 0]], np.float32)
    query_points = np.array([[1.1, 0.0, 0.0], [-1.2, 0.0, 0.0], [0.0, 1.3, 0.0], [0.0, -1.4, 0.0]], np.float32)
    pm = PMFTXY(3, 3, 3)
    pm.compute((box, points), np.zeros(4, np.float32), query_points,
               neighbors={"mode": "nearest", "num_neighbors": 1})
    npt.assert_array_equal(pm.bin_counts, [[0, 1, 0], [1, 0, 1], [0, 1, 0]])
    pm.compute((box, query_points), np.zeros(1, np.float32), points,
               neighbors={"mode": "nearest", "num_neighbors": 1})
    npt.assert_array_equal(pm.bin_counts, [[0, 0, 0], [0, 0, 0], [0, 1, 0]])
    # test_pmft.py:487-521: the query point's orientation applies, as quaternions too
    p1 = np.array([[0.1, 0.1, 0]], np.float32)
    p2 = np.array([[1, 0, 0]], np.float32)
    pm = PMFTXY(3, 3, 12)
    quat = np.array([[1, 0, 0, 0]], np.float32)
    pm.compute((box, p1), quat, p2, neighbors={"mode": "nearest", "num_neighbors": 1})
    assert tuple(np.asarray(np.where(pm.bin_counts)).squeeze()) == (4, 6)
    pm.compute((box, p2), np.zeros(1, np.float32), p1, neighbors={"mode": "nearest", "num_neighbors": 1})
    assert tuple(np.asarray(np.where(pm.bin_counts)).squeeze()) == (7, 5)
    with pytest.raises(ValueError, match="only supports"):
        pm.compute((Box.cube(8), p1), np.zeros(1, np.float32))


# ------------------------------------------------------------------ PMFTXYT, PMFTR12, PMFTXYZ
_TWO_PI = 2 * np.pi


def _quat_axis_angle(axis, angles):
    axis = np.asarray(axis, np.float64) / np.linalg.norm(axis)
    a = np.atleast_1d(np.asarray(angles, np.float64))
    return np.concatenate([np.cos(a / 2)[:, None], np.sin(a / 2)[:, None] * axis[None, :]], 1)


def _alternating_lattice(size, angle=0.0, extra_shell=2):
    # tests/util.py:66-112 of the reference
    c, s = np.cos(angle), np.sin(angle)
    rot = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])
    p1, p2 = [], []
    for i in range(-extra_shell, size + extra_shell):
        for j in range(-extra_shell, size + extra_shell):
            p = rot.dot(np.array([i, j, 0]))
            if (i + j) % 2 == 0:
                p2.append(p)
            elif 0 <= i < size and 0 <= j < size:
                p1.append(p)
    return np.array(p1, np.float32), np.array(p2, np.float32)


def test_pmftxyz_matches_oracle(oracle):
    box, pts = freud_amd.data.make_random_system(9, 3000, seed=7)
    pts = pts.astype(np.float32)
    rng = np.random.default_rng(8)
    q = rng.normal(size=(len(pts), 4))
    q = (q / np.linalg.norm(q, axis=1, keepdims=True)).astype(np.float32)
    equiv = np.array([[1, 0, 0, 0], [0, 0, 0, 1]], np.float32)
    lim = (1.6, 1.7, 1.8)
    pm = freud_amd.pmft.PMFTXYZ(*lim, (20, 22, 24))
    pm.compute((box, pts), q, equiv_orientations=equiv)
    ref_nl = oracle.query(_bt(box), pts, pts, dict(r_max=float(pm.r_max), exclude_ii=True), nthreads=8)
    counts = oracle.pmft3_accumulate("xyz", ref_nl, lim, (20, 22, 24), q_quat=q, equiv=equiv)
    npt.assert_array_equal(pm.bin_counts, counts)
    npt.assert_array_equal(pm._pcf, oracle.pmft3_reduce("xyz", _bt(box), counts, lim, len(pts), len(pts), 1, 2))
    pm.compute((box, pts), q, equiv_orientations=equiv, reset=False)
    npt.assert_array_equal(pm.bin_counts, 2 * counts)
    with pytest.raises(RuntimeError, match="equivalent orientations"):
        pm.compute((box, pts), q, equiv_orientations=equiv[:1], reset=False)
    assert pm.nbins == [20, 22, 24]
    npt.assert_allclose(pm.bin_edges[2], np.linspace(-1.8, 1.8, 25, dtype=np.float32), atol=1e-5)
    # a 2D box is refused (Box::enforce3D)
    with pytest.raises(ValueError, match="only supports"):
        freud_amd.pmft.PMFTXYZ(*lim, 4).compute((Box.square(9), pts * [1, 1, 0]), q)


def _angle_parity(gpu, ref, axes):
    """Counts equal except for bonds whose angle lies within 4 ulp of a bin
    edge of its angle axis (the reference's angle is the host library's float
    atan2, the GPU's atan2 in double rounded to float).  axes: [(angles of
    every bond in float64, bins of that axis)]."""
    risky = None
    for t_vals, n_t in axes:
        u = t_vals * n_t / _TWO_PI
        r = np.abs(u - np.round(u)) < 4 * np.spacing(np.float32(_TWO_PI)) * n_t / _TWO_PI
        risky = r if risky is None else risky | r
    diff = gpu.astype(np.int64) - ref.astype(np.int64)
    assert np.abs(diff).sum() <= 2 * int(risky.sum())
    assert abs(int(gpu.sum()) - int(ref.sum())) <= int(risky.sum())
    return int(np.abs(diff).sum()), int(risky.sum())


@pytest.mark.parametrize("cls", ["xyt", "r12"])
def test_pmft_angles_match_oracle(oracle, cls):
    box, pts = freud_amd.data.make_random_system(20, 6000, is2D=True, seed=11)
    pts = pts.astype(np.float32)
    rng = np.random.default_rng(12)
    theta = (rng.random(len(pts)) * 2 * np.pi).astype(np.float32)
    if cls == "xyt":
        pm = freud_amd.pmft.PMFTXYT(2.2, 2.3, (30, 32, 36))
        maxes, bins = (2.2, 2.3), (30, 32, 36)
    else:
        pm = freud_amd.pmft.PMFTR12(3.1, (25, 36, 40))
        maxes, bins = (3.1,), (25, 36, 40)
    pm.compute((box, pts), theta)
    ref_nl = oracle.query(_bt(box), pts, pts, dict(r_max=float(pm.r_max), exclude_ii=True), nthreads=8)
    counts = oracle.pmft3_accumulate(cls, ref_nl, maxes, bins, p_theta=theta, q_theta=theta)
    v = ref_nl["vectors"].astype(np.float64)
    i, j = ref_nl["ij"][:, 0], ref_nl["ij"][:, 1]
    if cls == "xyt":
        t = (theta[j] - np.arctan2(-v[:, 1], -v[:, 0])) % _TWO_PI
        d, r = _angle_parity(pm.bin_counts, counts, [(t, bins[2])])
    else:
        t1 = (theta[j] - np.arctan2(v[:, 1], v[:, 0])) % _TWO_PI
        t2 = (theta[i] - np.arctan2(-v[:, 1], -v[:, 0])) % _TWO_PI
        d, r = _angle_parity(pm.bin_counts, counts, [(t1, bins[1]), (t2, bins[2])])
    print(f"{cls}: {d} count differences, {r} bonds within 4 ulp of an angle edge", flush=True)
    pcf = oracle.pmft3_reduce(cls, _bt(box), pm.bin_counts, maxes, len(pts), len(pts), 1)
    npt.assert_array_equal(pm._pcf, pcf)


def test_pmft_angle_known_answers():
    # two particles (test_pmft.py:119-154)
    pts = np.array([[-1.0, 0.0, 0.0], [1.0, 0.1, 0.0]], np.float32)
    quats = np.array([[1, 0, 0, 0], [1, 0, 0, 0]], np.float32)
    pm = freud_amd.pmft.PMFTR12(5.23, (10, 20, 30))
    pm.compute((Box.square(16), pts), quats)
    exp = np.zeros((10, 20, 30), np.int64)
    for q, p in ((0, 1), (1, 0)):
        r = pts[p].astype(np.float64) - pts[q]
        exp[int(np.floor(np.linalg.norm(r) * 10 / 5.23)), int(np.floor((-np.arctan2(r[1], r[0]) % _TWO_PI) * 20 / _TWO_PI)),
            int(np.floor((-np.arctan2(-r[1], -r[0]) % _TWO_PI) * 30 / _TWO_PI))] += 1
    npt.assert_array_equal(pm.bin_counts, exp)
    pm = freud_amd.pmft.PMFTXYT(3.6, 4.2, (20, 30, 40))
    pm.compute((Box.square(16), pts), quats)
    lim = np.array([3.6, 4.2])
    exp = np.zeros((20, 30, 40), np.int64)
    for q, p in ((0, 1), (1, 0)):
        r = pts[p].astype(np.float64) - pts[q]
        xy = np.floor((r[:2] + lim) * np.array([20, 30]) / (2 * lim)).astype(int)
        exp[xy[0], xy[1], int(np.floor((-np.arctan2(-r[1], -r[0]) % _TWO_PI) * 40 / _TWO_PI))] += 1
    npt.assert_array_equal(pm.bin_counts, exp)
    with pytest.raises(ValueError, match="only supports"):
        pm.compute((Box.cube(16), pts), quats)
    # points != query points on an alternating lattice (test_pmft.py:248-270, 315-344)
    points, query_points = _alternating_lattice(10, 0.01, 2)
    box = Box.square(50)
    pm = freud_amd.pmft.PMFTR12(2.3, 10)
    pm.compute((box, points), np.zeros(len(points)), query_points, np.zeros(len(query_points)),
               neighbors=dict(mode="ball", r_max=2.3, exclude_ii=False))
    with np.errstate(divide="ignore"):
        assert np.count_nonzero(np.isinf(pm.pmft) == 0) == 12
        assert len(np.unique(pm.pmft)) == 3
    pm = freud_amd.pmft.PMFTXYT(2.5, 2.5, (10, 10, 4))
    pm.compute((box, points), np.zeros(len(points)), query_points, np.zeros(len(query_points)),
               neighbors=dict(mode="ball", r_max=float(np.sqrt(2 * 2.5**2)), exclude_ii=False))
    with np.errstate(divide="ignore"):
        for t in range(4):
            assert np.count_nonzero(np.isinf(pm.pmft[..., t]) == 0) == 3
        assert len(np.unique(pm.pmft)) == 2


@pytest.mark.parametrize("angle", [0.0, np.pi / 4])
@pytest.mark.parametrize("query_angle", [0.01, np.pi / 4 + 0.01])
def test_pmftxyt_nontrivial_orientations(angle, query_angle):
    # test_pmft.py:346-376
    p = np.array([[-1.0, 0.0, 0.0]], np.float32)
    q = np.array([[0.9, 0.1, 0.0]], np.float32)
    pm = freud_amd.pmft.PMFTXYT(2, 2, 4)
    pm.compute((Box.square(6), p), [angle], q, [query_angle])
    r = p[0].astype(np.float64) - q[0]
    c, s = np.cos(-query_angle), np.sin(-query_angle)
    rr = np.array([c * r[0] - s * r[1], s * r[0] + c * r[1]])
    xy = np.floor((rr + 2) * 4 / 4).astype(int)
    tb = int(np.floor(((angle - np.arctan2(-r[1], -r[0])) % _TWO_PI) * 4 / _TWO_PI))
    assert tuple(np.asarray(np.where(pm.bin_counts)).flatten()) == (xy[0], xy[1], tb)
    assert pm.bin_counts.sum() == 1


def test_pmftxyz_known_answers():
    # test_pmft.py:708-727 (shift), 729-779 (nearest neighbours)
    points = np.array([[1, 1, 1], [0, 0, 0]], np.float32)
    qo = np.array([[1, 0, 0, 0], [1, 0, 0, 0]], np.float32)
    noshift = freud_amd.pmft.PMFTXYZ(0.5, 0.5, 0.5, 3, shiftvec=[0, 0, 0])
    shift = freud_amd.pmft.PMFTXYZ(0.5, 0.5, 0.5, 3, shiftvec=[1, 1, 1])
    for pm in (noshift, shift):
        pm.compute((Box.cube(3), points), qo)
    with np.errstate(divide="ignore"):
        assert np.isfinite(noshift.pmft).sum() == 0
        assert np.isfinite(shift.pmft).sum() == 1
    box = Box.cube(8)
    pts = np.array([[0, 0, 0]], np.float32)
    qp = np.array([[1.1, 0, 0], [-1.2, 0, 0], [0, 1.3, 0], [0, -1.4, 0], [0, 0, 1.5], [0, 0, -1.6]], np.float32)
    pm = freud_amd.pmft.PMFTXYZ(3, 3, 3, 3)
    pm.compute((box, pts), np.tile([1.0, 0, 0, 0], (6, 1)), qp, neighbors={"mode": "nearest", "num_neighbors": 1})
    npt.assert_array_equal(pm.bin_counts, [[[0, 0, 0], [0, 1, 0], [0, 0, 0]], [[0, 1, 0], [1, 0, 1], [0, 1, 0]],
                                           [[0, 0, 0], [0, 1, 0], [0, 0, 0]]])
    pm.compute((box, qp), [[1.0, 0, 0, 0]], pts, neighbors={"mode": "nearest", "num_neighbors": 1})
    assert pm.bin_counts[2, 1, 1] == 1 and pm.bin_counts.sum() == 1
    # the query point's orientation applies (test_pmft.py:781-857)
    p1 = np.array([[0.1, 0.1, 0]], np.float32)
    p2 = np.array([[1, 0, 0]], np.float32)
    pm = freud_amd.pmft.PMFTXYZ(3, 3, 3, 12)
    pm.compute((box, p1), _quat_axis_angle([0, 0, 1], 0.0), p2, neighbors={"mode": "nearest", "num_neighbors": 1})
    npt.assert_array_equal(np.asarray(np.where(pm.bin_counts)).squeeze(), (4, 6, 6))
    pm.compute((box, p2), _quat_axis_angle([0, 0, 1], 0.0), p1, neighbors={"mode": "nearest", "num_neighbors": 1})
    npt.assert_array_equal(np.asarray(np.where(pm.bin_counts)).squeeze(), (7, 5, 6))


# ------------------------------------------------------------------ CorrelationFunction
def test_correlation_known_answers():
    # test_density_correlation_function.py:247-377 of the reference
    box = Box.cube(8)
    points = np.array([[0, 0, 0]], np.float32)
    qp = np.array([[0.4, 0, 0], [0.9, 0, 0], [0, 1.4, 0], [0, 1.9, 0], [0, 0, 2.4], [0, 0, 2.9]], np.float32)
    nn = {"mode": "nearest", "num_neighbors": 1}
    cf = freud_amd.density.CorrelationFunction(3, 3)
    with pytest.raises(AttributeError):
        cf.correlation
    cf.compute((box, points), np.ones(1), qp, np.ones(6), neighbors=nn)
    npt.assert_array_equal(cf.correlation, [1, 1, 1])
    npt.assert_array_equal(cf.bin_counts, [2, 2, 2])
    assert not np.iscomplexobj(cf.correlation)
    cf.compute((box, qp), np.ones(6), points, np.ones(1), neighbors=nn)
    npt.assert_array_equal(cf.correlation, [1, 0, 0])
    npt.assert_array_equal(cf.bin_counts, [1, 0, 0])
    values = [1 + 1j]
    qv = [1 + 1j, 1 + 1j, 2 + 2j, 2 + 2j, 3 + 3j, 3 + 3j]
    cf.compute((box, points), values, qp, qv, neighbors=nn)
    npt.assert_array_equal(cf.correlation, [2, 4, 6])
    cf.compute((box, points), np.conj(values), qp, qv, neighbors=nn)
    npt.assert_array_equal(cf.correlation, [2j, 4j, 6j])
    cf.compute((box, qp), np.conj(qv), points, values, neighbors=nn)
    npt.assert_array_equal(cf.correlation, [2j, 0, 0])
    npt.assert_array_equal(cf.bin_counts, [1, 0, 0])
    with pytest.raises(ValueError):
        freud_amd.density.CorrelationFunction(0, 3)
    with pytest.raises(ValueError):
        freud_amd.density.CorrelationFunction(3, -1)


@pytest.mark.parametrize("bins", [40, 3000])
def test_correlation_matches_oracle(oracle, bins):
    box, pts = freud_amd.data.make_random_system(31, 4000, is2D=True, seed=1)
    pts = pts.astype(np.float32)
    rng = np.random.default_rng(3)
    v = np.exp(1j * rng.random(len(pts)) * 2 * np.pi) * (1 + rng.random(len(pts)))
    cf = freud_amd.density.CorrelationFunction(bins, 10.0).compute((box, pts), v)
    ref_nl = oracle.query(_bt(box), pts, pts, dict(r_max=10.0, exclude_ii=True), nthreads=8)
    c, corr = oracle.correlation_accumulate(ref_nl, v, v, bins, 10.0)
    npt.assert_array_equal(cf.bin_counts, c)
    # complex double sums in another order than the oracle's bond order
    npt.assert_allclose(cf.correlation, corr, rtol=1e-11, atol=1e-13)
    assert np.iscomplexobj(cf.correlation)
    # reset=False adds a second frame with other values
    w = rng.random(len(pts)) - 0.5
    cf.compute((box, pts), w, reset=False)
    c2, _ = oracle.correlation_accumulate(ref_nl, w, w, bins, 10.0)
    npt.assert_array_equal(cf.bin_counts, c + c2)
    assert cf.box == box


# ------------------------------------------------------------------ BondOrder
def _rand_quats(n, seed):
    q = np.random.default_rng(seed).normal(size=(n, 4))
    return (q / np.linalg.norm(q, axis=1, keepdims=True)).astype(np.float32)


def test_bond_order_reference_cases():
    # test_environment_bond_order.py:14-140 of the reference
    box, pos = freud_amd.data.UnitCell.fcc().generate_system(4)
    quats = np.tile(np.array([1, 0, 0, 0], np.float32), (len(pos), 1))
    bo = freud_amd.environment.BondOrder(6)
    with pytest.raises(AttributeError):
        bo.bond_order
    bo.compute((box, pos), quats, neighbors={"num_neighbors": 12, "r_max": 1.5})
    op = bo.bond_order.copy()
    assert np.sum(op > 0) == 12
    assert bo.nbins == [6, 6] and bo.box == box
    npt.assert_allclose(bo.bin_centers[0], (2 * np.arange(6) + 1) * np.pi / 6, atol=1e-6)
    npt.assert_allclose(bo.bin_centers[1], (2 * np.arange(6) + 1) * np.pi / 12, atol=1e-6)
    nn = {"mode": "nearest", "num_neighbors": 12, "r_guess": 1.5, "exclude_ii": True}
    for mode in ("lbod", "obcd"):
        b = freud_amd.environment.BondOrder(6, mode=mode)
        b.compute((box, pos), quats, neighbors=nn, reset=False)
        npt.assert_allclose(b.bond_order, op)
    rq = _rand_quats(len(pos), 10893)
    npt.assert_allclose(freud_amd.environment.BondOrder(6).compute((box, pos), rq, neighbors=nn).bond_order, op)
    b = freud_amd.environment.BondOrder(6, mode="obcd").compute((box, pos), rq, neighbors=nn)
    assert not np.allclose(b.bond_order, op) and np.sum(b.bond_order > 0) == b.bond_order.size
    b = freud_amd.environment.BondOrder(6, mode="oocd").compute((box, pos), quats, neighbors=nn)
    assert np.sum(b.bond_order > 0) == 1 and b.bond_order[0, 0] > 0
    assert np.sum(freud_amd.environment.BondOrder(6, mode="oocd").compute(
        (box, pos), rq, neighbors=nn).bond_order > 0) > 30
    # points != query points
    query_points, points = _alternating_lattice(10, np.pi / 30)
    b = freud_amd.environment.BondOrder(bins=(35, 2))
    b.compute((Box.square(50), points), np.tile([1.0, 0, 0, 0], (len(points), 1)), query_points,
              np.tile([1.0, 0, 0, 0], (len(query_points), 1)),
              neighbors={"mode": "nearest", "num_neighbors": 12, "r_guess": 1.6, "exclude_ii": False})
    assert np.count_nonzero(b.bond_order) == 12
    assert len(np.unique(b.bond_order)) == 2
    with pytest.raises(NotImplementedError):
        freud_amd.environment.BondOrder(6).compute((box, pos), quats)
    with pytest.raises(ValueError):
        freud_amd.environment.BondOrder(6, mode="nope")
    with pytest.raises(ValueError):
        freud_amd.environment.BondOrder(1)


@pytest.mark.parametrize("mode", ["bod", "lbod", "obcd", "oocd"])
def test_bond_order_matches_oracle(oracle, mode):
    box, pts = freud_amd.data.make_random_system(12, 6000, seed=13)
    pts = pts.astype(np.float32)
    pq, qq = _rand_quats(len(pts), 1), _rand_quats(len(pts), 2)
    bo = freud_amd.environment.BondOrder((30, 18), mode=mode)
    qp = pts + np.float32(0.01)
    bo.compute((box, pts), pq, qp, qq, neighbors={"r_max": 1.6})
    ref_nl = oracle.query(_bt(box), pts, qp, dict(r_max=1.6), nthreads=8)
    counts, ref_bo = oracle.bond_order(ref_nl, pq, qq, mode, 30, 18)
    i, j = ref_nl["ij"][:, 0], ref_nl["ij"][:, 1]
    v = ref_nl["vectors"].astype(np.float64)

    def rot(q, x):
        q = np.asarray(q, np.float64)
        s, u = q[:, :1], q[:, 1:]
        return (s * s - np.sum(u * u, 1, keepdims=True)) * x + 2 * s * np.cross(u, x) + 2 * np.sum(
            u * x, 1, keepdims=True) * u

    pc = pq[j].astype(np.float64) * [1, -1, -1, -1]
    if mode == "lbod":
        v = rot(pc, v)
    elif mode == "obcd":
        v = rot(qq[i], rot(pc, v))
    elif mode == "oocd":
        v = rot(pc, rot(qq[i], np.tile([0.0, 0, 1], (len(i), 1))))
    theta = np.arctan2(v[:, 1], v[:, 0]) % _TWO_PI
    phi = np.arccos(v[:, 2] / np.linalg.norm(v, axis=1))
    d, r = _angle_parity(bo.bin_counts, counts, [(theta, 30), (2 * phi, 18)])
    print(f"{mode}: {d} count differences, {r} bonds within 4 ulp of an angle edge", flush=True)
    if d == 0:
        npt.assert_array_equal(bo.bond_order, ref_bo)
