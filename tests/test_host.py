"""Host-side checks that need no GPU: the C-ABI library loads and exports every
symbol of include/freud_amd.h, the host Box matches the oracle bit for bit, and
the Python argument handling mirrors the reference."""

import ctypes
import os
import re

import numpy as np
import numpy.testing as npt
import pytest

from conftest import ROOT
from freud_amd import _lib
from freud_amd.box import Box


def _header_symbols():
    txt = open(os.path.join(ROOT, "include", "freud_amd.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*|double)\s+(fa_\w+)\s*\(", txt, re.M)))


def test_library_exports_header_symbols():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    handle = ctypes.CDLL(_lib.LIB_PATH)
    syms = _header_symbols()
    assert len(syms) > 30
    missing = [s for s in syms if not hasattr(handle, s)]
    assert not missing, missing
    # every header symbol has a Python signature
    assert set(syms) <= set(_lib.exported_symbols())


def test_box_bit_exact_vs_oracle(oracle):
    rng = np.random.default_rng(0)
    for p in [(10, 12, 14, 0.3, -0.2, 0.1, False), (9, 9, 0, 0.2, 0, 0, True), (3, 3, 3, 0, 0, 0, False)]:
        box = Box(*p)
        v = ((rng.random((500, 3)) - 0.5) * 40).astype(np.float32)
        if box.is2D:
            v[:, 2] = 0
        npt.assert_array_equal(box.wrap(v), oracle.box_op(p, v, "wrap"))
        npt.assert_array_equal(box.make_fractional(v), oracle.box_op(p, v, "fractional"))
        f = rng.random((100, 3)).astype(np.float32)
        npt.assert_array_equal(box.make_absolute(f), oracle.box_op(p, f, "absolute"))
        npt.assert_array_equal(box.nearest_plane_distance(), oracle.nearest_plane_distance(p))


def test_box_api():
    box = Box(2, 2, 2, 1, 0, 0)
    npt.assert_allclose(box.wrap([[10, -5, -5], [0, 0.5, 0]]), [[-2, -1, -1], [0, 0.5, 0]], rtol=1e-6)
    assert Box.cube(3).volume == 27
    assert Box.square(3).is2D and Box.square(3).dimensions == 2
    assert Box.from_box([1, 2, 3]) == Box(1, 2, 3)
    assert Box.from_box({"Lx": 1, "Ly": 2}).is2D
    b = Box.from_matrix(Box(4, 5, 6, 0.1, 0.2, 0.3).to_matrix())
    npt.assert_allclose([b.Lx, b.Ly, b.Lz, b.xy, b.xz, b.yz], [4, 5, 6, 0.1, 0.2, 0.3], rtol=1e-6)
    with pytest.raises(ValueError):
        Box(0, 1, 1)
    img = Box.cube(2).get_images([[1.5, 0, -2.5]])
    npt.assert_array_equal(img, [[1, 0, -1]])


def test_query_args():
    from freud_amd.locality import _QueryArgs

    qa = _QueryArgs.from_dict(dict(r_max=3, exclude_ii=True))
    assert qa.mode is None and qa.r_max == 3 and qa.exclude_ii
    qa = _QueryArgs(mode="nearest", num_neighbors=4)
    assert qa.mode == "nearest" and qa.num_neighbors == 4
    with pytest.raises(ValueError):
        _QueryArgs(bad=1)
    with pytest.raises(ValueError):
        _QueryArgs(mode="sphere")


def test_rdf_arguments_and_axis():
    from freud_amd.density import RDF

    for args in [(0, 1.0), (10, -1.0), (10, 1.0, -0.5), (10, 1.0, 2.0)]:
        with pytest.raises(ValueError):
            RDF(*args)
    with pytest.raises(ValueError):
        RDF(10, 1.0, normalization_mode="bad")
    r = RDF(37, 3.3, 0.5)
    npt.assert_allclose(r.bin_edges, np.histogram_bin_edges([0.5, 3.3], bins=37, range=(0.5, 3.3)), rtol=1e-5,
                        atol=1e-6)
    assert r.nbins == 37 and r.bounds == pytest.approx((0.5, 3.3))
    with pytest.raises(AttributeError):
        r.rdf


def test_rdf_reduce_matches_oracle(oracle):
    # RDF.cc:73-99 normalisation in the reference's float order
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    rng = np.random.default_rng(0)
    for p, fs in [((10, 10, 10, 0, 0, 0, False), 0), ((8, 9, 0, 0, 0, 0, True), 1)]:
        counts = rng.integers(0, 5000, size=64).astype(np.uint32)
        box = _lib.fa_box(*p[:6], int(p[6]))
        rdf = np.empty(64, np.float32)
        n_r = np.empty(64, np.float32)
        rc = lib.fa_rdf_reduce(ctypes.byref(box), counts.ctypes.data_as(ctypes.c_void_p), ctypes.c_uint32(64),
                               ctypes.c_float(0.25), ctypes.c_float(4.0), ctypes.c_uint32(1000), ctypes.c_uint32(900),
                               ctypes.c_uint32(3), fs, rdf.ctypes.data_as(ctypes.c_void_p),
                               n_r.ctypes.data_as(ctypes.c_void_p))
        assert rc == 0
        ref_rdf, ref_nr = oracle.rdf_reduce(p, counts, 64, 0.25, 4.0, 1000, 900, 3, finite_size=bool(fs))
        npt.assert_array_equal(rdf, ref_rdf)
        npt.assert_array_equal(n_r, ref_nr)


def test_data_generators():
    from freud_amd import data

    box, pts = data.fcc_jittered(4, 1.5874, 0.0)
    assert pts.shape == (256, 3) and pts.dtype == np.float32
    f = box.make_fractional(pts)
    assert np.all((f >= 0) & (f < 1))
    box, pts = data.triclinic_uniform(1000, 0.85)
    assert np.all(np.abs(box.make_fractional(pts) - 0.5) <= 0.5)
    box, pts = data.hex2d_jittered(500)
    assert box.is2D and np.all(pts[:, 2] == 0)
    box, pts, lab = data.two_phase(3)
    assert len(pts) == 2 * 4 * 27 and lab.sum() == 4 * 27


def test_gsd_reader():
    from freud_amd.gsd import open_gsd

    from conftest import GOLDEN

    t = open_gsd(os.path.join(GOLDEN, "lj.gsd"))
    assert len(t) >= 2
    fr = t[0]
    assert fr.positions.shape == (1000, 3)
    npt.assert_allclose(fr.box[:3], 20.0)
    fr = open_gsd(os.path.join(GOLDEN, "steinhardt_average", "Test_Configuration.gsd"))[-1]
    assert fr.positions.shape == (3288, 3)


def test_dcd_reader():
    """lj.dcd and lj.gsd hold the same HOOMD run (tests/integration/files/lj/lj.py
    of the reference dumps both every 1000 steps): every DCD frame must equal
    the GSD frame bit for bit, with the 20 x 20 x 20 cubic box."""
    from freud_amd.dcd import open_dcd
    from freud_amd.gsd import open_gsd

    from conftest import GOLDEN

    d = open_dcd(os.path.join(GOLDEN, "lj.dcd"))
    g = open_gsd(os.path.join(GOLDEN, "lj.gsd"))
    assert len(d) == len(g) == 10 and d.n_atoms == 1000
    assert d.title[0] == "Created by HOOMD"
    for f in range(len(d)):
        a, b = d[f], g[f]
        assert a.step == b.step == 10000 + 1000 * f
        npt.assert_array_equal(a.positions, b.positions)
        assert a.cell == (20.0, 20.0, 20.0, 90.0, 90.0, 90.0)
        assert (a.box.Lx, a.box.Ly, a.box.Lz, a.box.xy, a.box.xz, a.box.yz, a.box.is2D) == (20, 20, 20, 0, 0, 0,
                                                                                            False)
    with pytest.raises(IndexError):
        d[10]
    # a triclinic cell from lengths and angles: |a| = A, |b| = B, |c| = C, and the angles between them
    from freud_amd.dcd import cell_vectors

    m = cell_vectors(10.0, 11.0, 12.0, 80.0, 85.0, 95.0)
    npt.assert_allclose(np.linalg.norm(m, axis=1), [10, 11, 12])
    cos = lambda u, v: np.dot(u, v) / np.linalg.norm(u) / np.linalg.norm(v)  # noqa: E731
    npt.assert_allclose(np.degrees(np.arccos([cos(m[1], m[2]), cos(m[0], m[2]), cos(m[0], m[1])])), [80, 85, 95])


def test_rdf_bin_table_is_exact(oracle):
    """The RDF kernels bin r_sq through a host-built table (thresholds + lookup);
    emulate the device arithmetic on many r_sq values and compare with the
    reference's RegularAxis bin of the correctly rounded sqrt."""
    rng = np.random.default_rng(0)
    for bins, r_min, r_max in ((200, 0.0, 5.0), (100, 0.0, 4.0), (37, 0.25, 3.9), (150, 0.5, 2.0), (7, 0.0, 1.0)):
        n = _lib.lib().fa_rdf_bin_table(bins, r_min, r_max, None, 0)
        assert n > bins
        tab = np.zeros(n, np.float32)
        assert _lib.lib().fa_rdf_bin_table(bins, r_min, r_max, tab.ctypes.data, n) == n
        T = tab[: bins + 1]
        inv_delta, x_fast = tab[bins + 1], tab[bins + 2]
        lut = tab[bins + 3:].view(np.uint16)[:4096]
        # r_sq spread over the range plus the thresholds themselves and their float neighbours
        x = np.concatenate([rng.uniform(0, 1.05 * float(T[-1]), 200000).astype(np.float32), T,
                            np.nextafter(T, np.float32(0)), np.nextafter(T, np.float32(np.inf))]).astype(np.float32)
        x = np.sort(x[(x >= T[0]) & (x < T[-1])])
        p = x * inv_delta
        assert p.max() < 4096  # the kernels index the table without a clamp
        c = p.astype(np.uint32)
        k = lut[c].astype(np.int64)
        ok = k != 0xFFFF
        assert np.all(ok[x >= x_fast])  # multi-bin cells lie below x_fast
        got = k[ok] + (x[ok] >= T[k[ok] + 1])
        # x is sorted and bins are monotone in x: equal histograms of two
        # non-decreasing bin sequences mean equal bins element by element
        assert np.all(np.diff(got) >= 0)
        ref = oracle.rdf_bin(np.sqrt(x[ok]), bins, r_min, r_max)  # histogram of the reference bins
        npt.assert_array_equal(np.bincount(got, minlength=bins), ref)
        assert ok.mean() > 0.9


def test_modulus_positive_shortcut_is_exact():
    """The kernels compute the reference's modulusPositive (utils.h:29-32),
    fmod(fmod(a, 1) + 1, 1), as t - floor(t) with t = a + 1 when |a| < 1;
    compare both on sampled float bit patterns and the edge values."""
    rng = np.random.default_rng(5)
    bits = rng.integers(0, 2 ** 32, 4_000_000, dtype=np.uint64).astype(np.uint32)
    a = bits.view(np.float32)
    one = np.float32(1.0)
    edge = np.array([0.0, -0.0, 1e-45, -1e-45, 2.0 ** -24, -(2.0 ** -24), np.nextafter(one, 0),
                     -np.nextafter(one, 0), 0.5, -0.5], np.float32)
    a = np.concatenate([a[np.abs(a) < 1], edge]).astype(np.float32)
    ref = np.fmod(np.fmod(a, one) + one, one)
    t = (a + one).astype(np.float32)
    fast = (t - np.floor(t)).astype(np.float32)
    npt.assert_array_equal(fast.view(np.uint32), ref.view(np.uint32))


def test_tuner_variant_counts():
    """The autotuner's variant counts per analysis (no GPU work): nine
    single-frame RDF cell shapes, five batched ones (keys offset by 1000),
    two batched ball-query cell sizes, four kNN cell sizes each with the ring
    or the staged search, and the two kNN searches on a reused grid."""
    assert _lib.tune_variants(1, 0) == 9
    assert _lib.tune_variants(1, 1003) == 5
    assert _lib.tune_variants(2, 1003) == 2
    assert _lib.tune_variants(3, 4) == 8
    assert _lib.tune_variants(4, 4) == 2
    with pytest.raises(ValueError):
        _lib.tune_variants(9, 0)


@pytest.mark.parametrize(
    ("box", "r_cut", "cells"),
    [
        # test_locality_neighbor_query.py:879-899 of the reference
        (Box.cube(10), 2.5, [7, 7, 7]),
        (Box.cube(10), 2.4999, [7, 7, 7]),
        (Box(1000, 100, 10), 1.0, [1003, 103, 13]),
        (Box(10000, 1000, 100), 1.0, [10003, 1003, 103]),
        (Box(2.0, 2.0, 2.0, 1.5, 1.0, 1.0), 1.0, [10, 7, 5]),
        (Box(1.0, 2.0, 3.0, 1.5, 0.5, 1.0), 0.4, [16, 15, 10]),
        (Box.square(5), 2.49, None),
        (Box(6.0, 6.0, 5.0, 0.1, -20.3, 0.0), 1.0, None),
    ],
)
def test_cellquery_grid_matches_oracle(oracle, box, r_cut, cells):
    """The library's CellQuery grid (host arithmetic) against the reference's
    known cell counts and, where the oracle can build the grid, its lower
    corner and inverse width bit for bit."""
    import brute

    n = np.zeros(3, np.int32)
    mn = np.zeros(3, np.float32)
    inv = ctypes.c_float()
    cbox = _lib.make_box(box)
    assert _lib.lib().fa_cellquery_grid(ctypes.byref(cbox), r_cut, n.ctypes.data, mn.ctypes.data,
                                        ctypes.byref(inv)) == 0
    if cells is not None:
        assert list(n) == cells
    assert tuple(n) == brute.cellquery_dims(box, r_cut)
    npt.assert_array_equal(mn, brute.cellquery_min_pos(box, r_cut))
    assert inv.value == np.float32(1) / np.float32(r_cut)
    if int(n[0]) * int(n[1]) * int(n[2]) < 10**7:
        g = oracle.cellquery_grid((box.Lx, box.Ly, box.Lz, box.xy, box.xz, box.yz, box.is2D), np.zeros((1, 3)),
                                  r_cut)
        assert tuple(n) == g["n"]
        npt.assert_array_equal(mn, g["min_pos"])
    # a non-positive width is refused like CellQuery::buildGrid
    assert _lib.lib().fa_cellquery_grid(ctypes.byref(cbox), 0.0, n.ctypes.data, mn.ctypes.data,
                                        ctypes.byref(inv)) != 0
