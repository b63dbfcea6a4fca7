"""Calls from several host threads at once.  Every C-ABI entry point holds the
library's lock while it runs (ctypes releases the GIL during the call), so
threads that compute RDFs, ball queries and Steinhardt order parameters on
different systems concurrently get the results of the same calls made one
after another."""

import threading

import numpy as np
import numpy.testing as npt
import pytest

import freud_amd
from freud_amd.box import Box

pytestmark = pytest.mark.gpu


def _system(seed):
    rng = np.random.default_rng(seed)
    L = 14.0 + seed % 3
    box = Box.cube(L) if seed % 2 == 0 else Box(L, L + 1.0, L + 2.0, 0.2, -0.1, 0.15)
    pts = box.make_absolute(rng.random((3000 + 500 * seed, 3)).astype(np.float32))
    return box, pts


def _work(box, pts):
    rdf = freud_amd.density.RDF(bins=80, r_max=4.0).compute((box, pts))
    nl = freud_amd.AABBQuery(box, pts).query(pts, dict(r_max=1.6, exclude_ii=True)).toNeighborList()
    st = freud_amd.order.Steinhardt(6).compute((box, pts), neighbors=dict(num_neighbors=12))
    return (np.array(rdf.bin_counts), np.array(nl[:]), np.array(nl.distances), np.array(st.particle_order))


def test_concurrent_threads_match_sequential():
    systems = [_system(s) for s in range(6)]
    expect = [_work(b, p) for b, p in systems]
    got = [None] * (2 * len(systems))
    errors = []

    def run(t):
        try:
            b, p = systems[t % len(systems)]
            for _ in range(3):
                got[t] = _work(b, p)
        except Exception as e:  # reported below, with the thread that raised it
            errors.append((t, repr(e)))

    threads = [threading.Thread(target=run, args=(t,)) for t in range(len(got))]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=120)
    assert not any(th.is_alive() for th in threads)
    assert not errors, errors
    for t, res in enumerate(got):
        ref = expect[t % len(systems)]
        for a, b in zip(res, ref):
            npt.assert_array_equal(a, b)
