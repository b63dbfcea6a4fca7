"""Benchmark of the pair-analysis hot path on one MI355X.

Workload (one "step" = one frame): N = 1,048,576 points, FCC 64^3 x 4 cells,
lattice constant a = 1.5874 (density 1.0), Gaussian jitter sigma = 0.05 a,
seed 1, periodic cube; per frame:
  * AABBQuery ball query r_max = 1.5 -> NeighborList   (grid built per frame)
  * kNN query k = 12 -> NeighborList
  * RDF, 200 bins, r_max = 5 (fused query + histogram)
  * Steinhardt l = 6: q6 + w6, and averaged q6, over the ball NeighborList
Inputs are resident in HBM before the timed region.  The metric is pairs/s:
(ball-query bonds + RDF pairs) per second of frame time.

Prints one JSON line (see the "roofline" and "cpu_baseline" objects).
"""

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_VALU_PEAK_TFLOPS = 157.3


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def make_workload():
    from freud_amd import data

    box, pts = data.fcc_jittered(64, 1.5874, 0.05 * 1.5874, seed=1)
    return box, pts


class Frame:
    """One frame of the pipeline through the C-ABI, with persistent outputs."""

    def __init__(self, box, d_pts, n):
        from freud_amd import _lib
        from freud_amd._lib import DeviceArray

        self.L = _lib
        self.lib = _lib.lib()
        self.box = box
        self.cbox = _lib.make_box(box)
        self.d_pts = d_pts
        self.n = n
        self.nl_ball = _lib.fa_nlist()
        self.nl_knn = _lib.fa_nlist()
        self.counts = DeviceArray.zeros(200, np.uint32)
        self.d_qlm = DeviceArray((n * 13 * 2,), np.float32)
        self.d_qlmave = DeviceArray((n * 13 * 2,), np.float32)
        self.d_ql = DeviceArray((n,), np.float32)
        self.d_qlave = DeviceArray((n,), np.float32)
        self.d_wl = DeviceArray((n,), np.float32)
        self.ls = np.array([6], np.uint32)
        self.order = np.zeros(1, np.float32)
        self.ball = _lib.fa_query_args(_lib.MODE_BALL, 0xFFFFFFFF, 1.5, 0.0, -1.0, -1.0, 1)
        self.knn = _lib.fa_query_args(_lib.MODE_NEAREST, 12, -1.0, 0.0, -1.0, -1.0, 1)

    def run(self):
        L, lib, check = self.L, self.lib, self.L.check
        h = ctypes.c_void_p()
        check(lib.fa_nq_create(ctypes.byref(self.cbox), self.d_pts.ptr, self.n, L.NQ_AABB, 0.0, ctypes.byref(h)))
        try:
            check(lib.fa_query(h, self.d_pts.ptr, self.n, ctypes.byref(self.ball), 0, ctypes.byref(self.nl_ball)))
            check(lib.fa_query(h, self.d_pts.ptr, self.n, ctypes.byref(self.knn), 0, ctypes.byref(self.nl_knn)))
            self.counts.zero()
            check(lib.fa_rdf_accumulate_query(h, self.d_pts.ptr, self.n, 5.0, 0.0, 1, 200, 0.0, 5.0,
                                              self.counts.ptr))
            check(lib.fa_steinhardt(ctypes.byref(self.cbox), self.d_pts.ptr, self.n, ctypes.byref(self.nl_ball),
                                    self.ls.ctypes.data, 1, L.ST_WL, self.d_qlm.ptr, self.d_ql.ptr, None, None,
                                    self.d_wl.ptr, self.order.ctypes.data))
            check(lib.fa_steinhardt(ctypes.byref(self.cbox), self.d_pts.ptr, self.n, ctypes.byref(self.nl_ball),
                                    self.ls.ctypes.data, 1, L.ST_AVERAGE, self.d_qlm.ptr, self.d_ql.ptr,
                                    self.d_qlave.ptr, self.d_qlmave.ptr, None, self.order.ctypes.data))
        finally:
            lib.fa_nq_destroy(h)

    def pairs(self):
        return int(self.nl_ball.num_bonds), int(self.counts.numpy().astype(np.int64).sum()), int(
            self.nl_knn.num_bonds)


def algorithmic_bytes(name, n, b_ball, b_knn, k=12, l=6, bins=200):
    """Compulsory HBM bytes per launch of each kernel (inputs read once,
    outputs written once).  The averaged q_lm also gathers 8(2l+1) bytes of
    neighbour q_lm per bond, mostly from cache; those re-reads are not
    compulsory and are not counted here."""
    nm = 2 * l + 1
    table = {
        # one walk: cell-ordered points in, counts + positions + packed keys out
        "ball_walk_kernel": 16 * n + 4 * n + 4 * n + 4 * b_ball,
        # emission: offsets, query and neighbour positions, positions, keys in; bonds out
        "ball_emit_kernel": 4 * (n + 1) + 12 * n + 12 * n + 4 * n + 4 * b_ball + 28 * b_ball,
        "knn_aabb_kernel": 12 * n + 12 * n + 4 * n + 4 * n + 8 * k * n,
        # staged search: cell-ordered points in; counts, positions, selections out
        "knn_stage_kernel": 16 * n + 4 * n + 4 * n + 8 * k * n,
        "knn_write_kernel": 8 * k * n + 4 * n + 4 * (n + 1) + 12 * n + 28 * b_knn,
        "rdf_self_kernel": 16 * n + 4 * bins,
        "rdf_cluster_kernel": 16 * n + 4 * bins,
        "qlm_kernel": 24 * b_ball + 8 * nm * n + 4 * n,
        "qlm_ave_kernel": 4 * b_ball + 4 * (n + 1) + 8 * nm * n + 8 * nm * n + 4 * n,
        "wl_kernel": 8 * nm * n + 4 * n + 4 * n,
        "grid_assign_kernel": 12 * n + 4 * n,
        "grid_scatter_kernel": 12 * n + 4 * n + 16 * n + 4 * n,
    }
    for key, v in table.items():
        if name.startswith(key) or key in name:
            return v
    return None


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE, scaled by the
    same counters on a copy of known size) from the committed PMC summary, or
    None when no summary covers the kernel."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_frame_1M.json")
    try:
        with open(path) as fh:
            table = json.load(fh).get("kernels", {})
    except (OSError, ValueError):
        return None
    name = kernel.strip("()")
    rec = table.get(name)
    if rec is None:
        # the library names a template family once (qlm_kernel<L>); rocprof names the instance
        base = name.split("<")[0]
        hits = [v for k, v in table.items() if k.split("<")[0] == base]
        rec = hits[0] if len(hits) == 1 else None
    return None if rec is None else int(rec["fetch_bytes"] + rec["write_bytes"])


def rdf_stencil_evaluations(frames, r_max):
    """The RDF work unit: candidate-pair evaluations of a full stencil of
    r_max/2-wide cells (5^3 cells in 3D, 5^2 in 2D) around every point,
    N * 125 * rho * (r_max/2)^3 per 3D frame, summed over frames."""
    total = 0.0
    for box, n in frames:
        if box.is2D:
            total += n * 25.0 * (n / (box.Lx * box.Ly)) * (r_max / 2.0) ** 2
        else:
            total += n * 125.0 * (n / box.volume) * (r_max / 2.0) ** 3
    return total


def rdf_roofline(kernel, avg_ms, evaluations, binned_pairs, traffic=None):
    """RDF against the FP32 vector peak: each candidate evaluation is 8 flops
    (3 sub, 3 mul, 2 add for r_sq); the binned ordered pairs are reported
    beside it."""
    achieved = 8.0 * evaluations / (avg_ms * 1e-3) / 1e12
    return {"bound": "valu", "achieved": round(achieved, 3), "peak": FP32_VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / FP32_VALU_PEAK_TFLOPS, 5), "traffic": traffic, "kernel": kernel,
            "avg_ms": round(avg_ms, 4),
            "work": "candidate-pair evaluations of the 5^3 stencil of r_max/2 cells, 8 flop each",
            "units_per_launch": round(evaluations), "flop_per_unit": 8,
            "binned_pairs_per_launch": int(binned_pairs),
            "binned_pairs_per_s": round(binned_pairs / (avg_ms * 1e-3))}


def qlm_flops_per_bond(ls):
    """Flops of one bond in the q_lm pass for the list of l: the bond direction
    (14: 3 sub, div, 3 for sin theta, 3 for rho^2, 2 for 1/rho, 2 mul), the
    associated-Legendre recurrence up to max(l) (8 per m >= 1 for P(m,m) and
    e^{i m phi}, 2 for P(m+1,m), 4 for each P(l,m) with l >= m+2) and the
    accumulation of w Y_lm for every requested (l, m >= 0) (4: 2 mul, 2 add)."""
    lmax = max(ls)
    f = 14 + 8 * lmax
    for m in range(lmax + 1):
        f += 2 if m + 1 <= lmax else 0
        f += 4 * max(0, lmax - m - 1)
    f += 4 * sum(l + 1 for l in ls)
    return f


def qlm_roofline(kernel, avg_ms, bonds, n, ls):
    """Fused q_lm pass against the FP32 vector peak; the compulsory HBM bytes
    (bond index + distance + neighbour position 24 B per bond, q_lm and q_l
    out) are reported beside it."""
    fpb = qlm_flops_per_bond(ls)
    achieved = fpb * bonds / (avg_ms * 1e-3) / 1e12
    hbm = 24 * bonds + 8 * n * sum(2 * l + 1 for l in ls) + 4 * n * len(ls)
    return {"bound": "valu", "achieved": round(achieved, 3), "peak": FP32_VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / FP32_VALU_PEAK_TFLOPS, 5), "traffic": None, "kernel": kernel,
            "avg_ms": round(avg_ms, 4), "work": f"bonds x {fpb} flop (l = {list(ls)})",
            "units_per_launch": int(bonds), "flop_per_unit": fpb,
            "compulsory_hbm_gbs": round(hbm / (avg_ms * 1e-3) / 1e9, 1)}


def _oracle_frame(oracle, bt, pts, q, cores):
    """The frame's pipeline on the oracle for query points q."""
    nl = oracle.query(bt, pts, q, dict(r_max=1.5, exclude_ii=True), nthreads=cores)
    knn = oracle.query(bt, pts, q, dict(num_neighbors=12, exclude_ii=True), nthreads=cores)
    counts = oracle.rdf_accumulate(bt, pts, q, 200, 5.0, exclude_ii=True, nthreads=cores)
    st_w = oracle.steinhardt(bt, pts, nl, [6], wl=True, nthreads=cores)
    st_a = oracle.steinhardt(bt, pts, nl, [6], average=True, nthreads=cores)
    return nl, knn, counts, st_w, st_a


def cpu_baseline(box, pts, budget_s=15.0):
    """The oracle (C++ restatement of the reference's AABB tree, thread-local
    histograms and Steinhardt loops) on the frame: on all host cores over the
    whole frame when it finishes within the budget (else over a prefix of the
    query points), and on one thread over a prefix of the query points.
    Returns the baseline record and, when the whole frame ran, the oracle's
    outputs (the checker of the GPU frame)."""
    import oracle

    oracle.build()
    n = len(pts)
    cores = os.cpu_count() or 1
    bt = (box.Lx, box.Ly, box.Lz, box.xy, box.xz, box.yz, box.is2D)

    def timed(threads, budget):
        s = 2048
        while True:
            q = pts[:s]
            t0 = time.perf_counter()
            res = _oracle_frame(oracle, bt, pts, q, threads)
            elapsed = time.perf_counter() - t0
            if elapsed > budget / 4 or s >= n:
                return s, elapsed, res
            s = min(s * 2, n)

    s, elapsed, res = timed(cores, budget_s)
    pairs = len(res[0]["ij"]) + int(res[2].astype(np.int64).sum())
    s1, elapsed1, res1 = timed(1, budget_s)
    pairs1 = len(res1[0]["ij"]) + int(res1[2].astype(np.int64).sum())
    rec = {
        "value": pairs / elapsed,
        "unit": "pairs/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"{s} of {n} query points of the same frame (all {n} points searched), full pipeline "
                   f"(AABB tree build + ball r=1.5 + kNN k=12 + RDF r=5 + q6/w6/avg-q6) in {elapsed:.2f} s on "
                   f"{cores} threads of {_cpu_model()}"),
        "single_thread": {"value": pairs1 / elapsed1, "unit": "pairs/s", "cores": 1,
                          "sample": f"{s1} of {n} query points, same pipeline, {elapsed1:.2f} s on 1 thread"},
    }
    return rec, (res if s == n else None)


def frame_parity(fr, box, pts, ref):
    """The timed frame's GPU outputs against the oracle's run of the same
    frame: NeighborLists and RDF counts bit for bit, q6 / w6 / averaged q6 as
    max |error| against the oracle's float32 and float64 evaluations."""
    import oracle
    from freud_amd._lib import copy_d2h

    nl, knn, counts, st_w, st_a = ref
    cores = os.cpu_count() or 1
    bt = (box.Lx, box.Ly, box.Lz, box.xy, box.xz, box.yz, box.is2D)
    n = len(pts)

    def same_nlist(c, r):
        b = int(c.num_bonds)
        if b != len(r["ij"]):
            return False
        return bool(np.array_equal(copy_d2h(c.neighbors, (b, 2), np.uint32), r["ij"])
                    and np.array_equal(copy_d2h(c.distances, (b,), np.float32), r["distances"])
                    and np.array_equal(copy_d2h(c.vectors, (b, 3), np.float32), r["vectors"]))

    ql = fr.d_ql.numpy().astype(np.float64)
    qlave = fr.d_qlave.numpy().astype(np.float64)
    wl = fr.d_wl.numpy().astype(np.float64)
    w64 = oracle.steinhardt(bt, pts, nl, [6], wl=True, f64=True, nthreads=cores)
    a64 = oracle.steinhardt(bt, pts, nl, [6], average=True, f64=True, nthreads=cores)

    def err(a, b):
        return float(np.max(np.abs(a - b)))

    out = {
        "ball_nlist_bit_exact": same_nlist(fr.nl_ball, nl),
        "knn_nlist_bit_exact": same_nlist(fr.nl_knn, knn),
        "rdf_counts_bit_exact": bool(np.array_equal(fr.counts.numpy(), counts)),
        "q6_max_err_vs_f32": err(ql, st_a["ql"][:, 0]),
        "q6_max_err_vs_f64": err(ql, a64["ql"][:, 0]),
        "w6_max_err_vs_f32": err(wl, st_w["wl"][:, 0]),
        "w6_max_err_vs_f64": err(wl, w64["wl"][:, 0]),
        "avg_q6_max_err_vs_f32": err(qlave, st_a["ql_ave"][:, 0]),
        "avg_q6_max_err_vs_f64": err(qlave, a64["ql_ave"][:, 0]),
        "oracle_f32_q6_err_vs_f64": err(st_a["ql"][:, 0], a64["ql"][:, 0]),
        "n_points": n,
    }
    out["pass"] = bool(out["ball_nlist_bit_exact"] and out["knn_nlist_bit_exact"] and out["rdf_counts_bit_exact"]
                       and max(out["q6_max_err_vs_f64"], out["w6_max_err_vs_f64"],
                               out["avg_q6_max_err_vs_f64"]) <= 1e-5)
    return out


class FrameLoop:
    """The per-frame plain GPU loop over a trajectory (what a freud user writes:
    one query / RDF / Steinhardt call sequence per frame), through the C-ABI."""

    def __init__(self, frames, bins=200, l=6):
        from freud_amd import _lib
        from freud_amd._lib import DeviceArray

        self.L = _lib
        self.lib = _lib.lib()
        self.frames = []
        for box, pts in frames:
            n = len(pts)
            self.frames.append((_lib.make_box(box), DeviceArray.from_numpy(pts), n))
        nmax = max(n for _, _, n in self.frames)
        self.nl = _lib.fa_nlist()
        self.counts = DeviceArray.zeros(len(frames) * bins, np.uint32)
        self.bins = bins
        self.d_qlm = DeviceArray((nmax * (2 * l + 1) * 2,), np.float32)
        self.d_ql = DeviceArray((nmax,), np.float32)
        self.ls = np.array([l], np.uint32)
        self.order = np.zeros(1, np.float32)
        self.ball = _lib.fa_query_args(_lib.MODE_BALL, 0xFFFFFFFF, 1.5, 0.0, -1.0, -1.0, 1)

    def run(self):
        L, lib, check = self.L, self.lib, self.L.check
        self.counts.zero()
        for f, (cbox, d_pts, n) in enumerate(self.frames):
            h = ctypes.c_void_p()
            check(lib.fa_nq_create(ctypes.byref(cbox), d_pts.ptr, n, L.NQ_AABB, 0.0, ctypes.byref(h)))
            try:
                check(lib.fa_query(h, d_pts.ptr, n, ctypes.byref(self.ball), 0, ctypes.byref(self.nl)))
                check(lib.fa_rdf_accumulate_query(h, d_pts.ptr, n, 5.0, 0.0, 1, self.bins, 0.0, 5.0,
                                                  self.counts.ptr + 4 * f * self.bins))
                check(lib.fa_steinhardt(ctypes.byref(cbox), d_pts.ptr, n, ctypes.byref(self.nl), self.ls.ctypes.data,
                                        1, 0, self.d_qlm.ptr, self.d_ql.ptr, None, None, None,
                                        self.order.ctypes.data))
            finally:
                lib.fa_nq_destroy(h)
        self.L.synchronize()


def run_batched(args):
    """256 frames, N log-uniform in [10k, 200k], liquid or jittered FCC/BCC,
    cubic or triclinic: ball query r=1.5 -> NeighborLists, RDF 200 bins r=5,
    Steinhardt q6 per frame.  One step = the whole trajectory, run as one
    batch replayed from a HIP graph; the per-frame loop is timed beside it."""
    from freud_amd import _lib, data
    from freud_amd.batch import TrajectoryBatch

    frames, kinds = data.mixed_frames(256, 10_000, 200_000, seed=4)
    total = sum(len(p) for _, p in frames)
    print(f"[bench] batched: {len(frames)} frames, {total} points", file=sys.stderr, flush=True)
    batch = TrajectoryBatch(frames, ball_r_max=1.5, rdf_bins=200, rdf_r_max=5.0, l=6)
    for w in range(max(args.warmup, 2)):
        batch.compute(use_graph=True)
        print(f"[bench] batch warmup {w} done, config {batch.config}", file=sys.stderr, flush=True)
    _lib.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.compute(use_graph=True)
    t1 = time.perf_counter()
    ms_batch = (t1 - t0) * 1e3 / args.steps
    bonds = int(batch.num_bonds.astype(np.int64).sum())
    counts = batch.rdf_counts
    rdf_pairs = int(counts.astype(np.int64).sum())
    # eager (no graph) batch for the per-kernel breakdown
    _lib.profile_reset()
    _lib.profile_enable(True)
    t2 = time.perf_counter()
    batch.compute(use_graph=False)
    t3 = time.perf_counter()
    _lib.profile_enable(False)
    stats = _lib.profile_stats()
    ms_eager = (t3 - t2) * 1e3
    print(f"[bench] batch {ms_batch:.2f} ms/trajectory (graph), {ms_eager:.2f} ms eager; bonds {bonds}, rdf pairs "
          f"{rdf_pairs}", file=sys.stderr, flush=True)

    loop = FrameLoop(frames)
    loop.run()
    steps_loop = max(1, min(args.steps, 3))
    t4 = time.perf_counter()
    for _ in range(steps_loop):
        loop.run()
    t5 = time.perf_counter()
    ms_loop = (t5 - t4) * 1e3 / steps_loop
    loop_counts = loop.counts.numpy().reshape(len(frames), 200)
    same = bool(np.array_equal(loop_counts, counts))
    print(f"[bench] per-frame loop {ms_loop:.2f} ms/trajectory; batch/loop RDF counts identical: {same}",
          file=sys.stderr, flush=True)

    kernels = {}
    for name, (cnt, tot) in sorted(stats.items(), key=lambda kv: -kv[1][1]):
        kernels[name] = {"launches_per_step": cnt, "avg_ms": round(tot / cnt, 4), "share": round(tot / ms_eager, 4)}
    dom = max(stats.items(), key=lambda kv: kv[1][1])[0]
    cnt, tot = stats[dom]
    evaluations = rdf_stencil_evaluations([(b, len(p)) for b, p in frames], 5.0)
    out = {
        "metric": "pairs/s (query+RDF) & particles/s (SolidLiquid), plain vs batched/learned",
        "value": (bonds + rdf_pairs) / (ms_batch / 1e3),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": max(args.warmup, 2),
        "ms_per_step": ms_batch,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded liquid / jittered FCC / BCC frames)",
        "config": {
            "workload": "batched trajectory: 256 frames, N log-uniform in [10k, 200k], liquid (density 0.8-1.0) or "
                        "jittered FCC/BCC (sigma 0.03-0.1a), cubic or triclinic (tilts in [-0.3, 0.3]); ball query "
                        "r_max=1.5 + RDF 200 bins r_max=5 + Steinhardt q6, one batch per step, HIP-graph replay",
            "n_frames": len(frames),
            "n_points": total,
            "ball_bonds": bonds,
            "rdf_pairs": rdf_pairs,
            "particles_per_s": total / (ms_batch / 1e3),
            "batch_config": batch.config,
            "ms_eager_batch": ms_eager,
            "ms_per_frame_loop": ms_loop,
            "speedup_vs_per_frame_loop": ms_loop / ms_batch,
            "loop_counts_identical": same,
        },
        "roofline": rdf_roofline(dom, tot / cnt, evaluations, rdf_pairs),
        "kernels": kernels,
        "cpu_baseline": None,
    }
    return _emit(out, args)


def run_stress(args):
    """Triclinic N=500,000 at density 0.85 (tilts U[-0.3, 0.3], seed 2): ball
    query r_max=2 -> NeighborList, RDF 150 bins r_max=2, Steinhardt q_l for
    l = 4, 6, 8, 10, 12; plus a 2D hexagonal lattice of 250,000 points (jitter
    0.05): kNN k=6 -> Hexatic k=6, and a ball query r_max=2."""
    from freud_amd import _lib, data
    from freud_amd._lib import DeviceArray

    L, lib, check = _lib, _lib.lib(), _lib.check
    tbox, tpts = data.triclinic_uniform(500_000, 0.85, seed=2, tilt=0.3)
    hbox, hpts = data.hex2d_jittered(250_000, jitter=0.05, seed=3)
    nt, nh = len(tpts), len(hpts)
    d_t, d_h = DeviceArray.from_numpy(tpts), DeviceArray.from_numpy(hpts)
    cb_t, cb_h = L.make_box(tbox), L.make_box(hbox)
    ls = np.array([4, 6, 8, 10, 12], np.uint32)
    nm = int(np.sum(2 * ls + 1))
    d_qlm = DeviceArray((nt * nm * 2,), np.float32)
    d_ql = DeviceArray((nt * len(ls),), np.float32)
    d_psi = DeviceArray((nh * 2,), np.float32)
    counts = DeviceArray.zeros(150, np.uint32)
    order = np.zeros(len(ls), np.float32)
    nl_t, nl_h, nl_k = L.fa_nlist(), L.fa_nlist(), L.fa_nlist()
    ball2 = L.fa_query_args(L.MODE_BALL, 0xFFFFFFFF, 2.0, 0.0, -1.0, -1.0, 1)
    knn6 = L.fa_query_args(L.MODE_NEAREST, 6, -1.0, 0.0, -1.0, -1.0, 1)

    def step():
        h = ctypes.c_void_p()
        check(lib.fa_nq_create(ctypes.byref(cb_t), d_t.ptr, nt, L.NQ_AABB, 0.0, ctypes.byref(h)))
        try:
            check(lib.fa_query(h, d_t.ptr, nt, ctypes.byref(ball2), 0, ctypes.byref(nl_t)))
            counts.zero()
            check(lib.fa_rdf_accumulate_query(h, d_t.ptr, nt, 2.0, 0.0, 1, 150, 0.0, 2.0, counts.ptr))
            check(lib.fa_steinhardt(ctypes.byref(cb_t), d_t.ptr, nt, ctypes.byref(nl_t), ls.ctypes.data, len(ls), 0,
                                    d_qlm.ptr, d_ql.ptr, None, None, None, order.ctypes.data))
        finally:
            lib.fa_nq_destroy(h)
        check(lib.fa_nq_create(ctypes.byref(cb_h), d_h.ptr, nh, L.NQ_AABB, 0.0, ctypes.byref(h)))
        try:
            check(lib.fa_query(h, d_h.ptr, nh, ctypes.byref(knn6), 0, ctypes.byref(nl_k)))
            check(lib.fa_hexatic(ctypes.byref(cb_h), d_h.ptr, nh, ctypes.byref(nl_k), 6, 0, d_psi.ptr))
            check(lib.fa_query(h, d_h.ptr, nh, ctypes.byref(ball2), 0, ctypes.byref(nl_h)))
        finally:
            lib.fa_nq_destroy(h)

    for w in range(args.warmup):
        step()
        L.synchronize()
    L.profile_reset()
    L.profile_enable(True)
    L.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    L.synchronize()
    t1 = time.perf_counter()
    L.profile_enable(False)
    stats = L.profile_stats()
    ms = (t1 - t0) * 1e3 / args.steps
    bt, bh, bk = int(nl_t.num_bonds), int(nl_h.num_bonds), int(nl_k.num_bonds)
    rdf_pairs = int(counts.numpy().astype(np.int64).sum())
    pairs = bt + bh + rdf_pairs
    print(f"[bench] stress {ms:.3f} ms/step: triclinic bonds {bt}, rdf pairs {rdf_pairs}, 2D ball bonds {bh}, "
          f"2D kNN bonds {bk}", file=sys.stderr, flush=True)
    kernels = {}
    for name, (cnt, tot) in sorted(stats.items(), key=lambda kv: -kv[1][1]):
        kernels[name] = {"launches_per_step": cnt / args.steps, "avg_ms": round(tot / cnt, 4),
                         "share": round(tot / (ms * args.steps), 4)}
    dom = max(stats.items(), key=lambda kv: kv[1][1])[0]
    cnt, tot = stats[dom]
    avg_ms = tot / cnt
    # dominant kernel: the fused q_lm pass against the FP32 vector peak, any
    # other against HBM with its compulsory bytes
    if "qlm_even_multi" in dom:
        roof = qlm_roofline(dom, avg_ms, bt, nt, [int(v) for v in ls])
    else:
        ab = algorithmic_bytes(dom, nt, bt, 0) or 0
        achieved = ab / (avg_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": dom, "avg_ms": round(avg_ms, 4)}
    out = {
        "metric": "pairs/s (query+RDF) & particles/s (SolidLiquid), plain vs batched/learned",
        "value": pairs / (ms / 1e3),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded uniform triclinic + jittered hexagonal lattice)",
        "config": {
            "workload": "stress: triclinic N=500000 density 0.85 tilts U[-0.3,0.3] (ball r_max=2, RDF 150 bins "
                        "r_max=2, Steinhardt l=4,6,8,10,12) + 2D hexagonal N=250000 jitter 0.05 (kNN k=6 + Hexatic "
                        "k=6, ball r_max=2)",
            "n_points": nt + nh,
            "triclinic_bonds": bt,
            "rdf_pairs": rdf_pairs,
            "hex_ball_bonds": bh,
            "hex_knn_bonds": bk,
            "particles_per_s": (nt + nh) / (ms / 1e3),
        },
        "roofline": roof,
        "kernels": kernels,
        "cpu_baseline": None,
    }
    return _emit(out, args)


def run_learned(args):
    """Two-phase box of 2,097,152 particles (FCC 64^3 x 4 at density 1.05,
    sigma 0.04 a, beside as many uniform-liquid particles at density 0.95):
    SolidLiquid l=6, q_threshold 0.7, solid_threshold 6 on the r_max=1.5
    NeighborList, exactly and through the learned classifier (trained on 32
    generated two-phase frames + 2 validation frames).  One step = the
    solid-like labels of the frame."""
    from freud_amd import _lib, data
    from freud_amd.learned import LearnedSolidLiquid
    from freud_amd.locality import AABBQuery
    from freud_amd.order import SolidLiquid

    t0 = time.perf_counter()
    frames = []
    for s in range(34):
        b, p, _ = data.two_phase(20, sigma=0.03 + 0.005 * (s % 5), seed=1000 + s)
        frames.append((b, p))
    model = LearnedSolidLiquid(q_threshold=0.7, solid_threshold=6, margin=0.5, seed=11)
    losses = model.fit(frames, epochs=6, batch=8192, lr=3e-3, validation=2, target_agreement=0.997)
    _lib.synchronize()
    t_train = time.perf_counter() - t0
    print(f"[bench] trained on {model._train_size} particles in {t_train:.1f} s; loss {losses[0]:.4f} -> "
          f"{losses[-1]:.4f}; margin {model.margin:.3f}", file=sys.stderr, flush=True)

    box, pts, _ = data.two_phase(64, rho_solid=1.05, rho_liquid=0.95, sigma=0.04, seed=5)
    n = len(pts)
    aq = AABBQuery(box, pts)
    nl = aq.query(pts, dict(r_max=1.5, exclude_ii=True)).toNeighborList()
    exact_sl = SolidLiquid(6, 0.7, 6)

    def run_exact():
        exact_sl.compute(aq, neighbors=nl)

    def run_learned():
        model.compute(aq, neighbors=nl)

    for _ in range(max(args.warmup, 1)):
        run_exact()
        run_learned()
    _lib.synchronize()
    exact = exact_sl.num_connections >= 6
    res = {}
    for name, fn in (("exact", run_exact), ("learned", run_learned)):
        _lib.profile_reset()
        _lib.profile_enable(True)
        _lib.synchronize()
        a = time.perf_counter()
        for _ in range(args.steps):
            fn()
        _lib.synchronize()
        b = time.perf_counter()
        _lib.profile_enable(False)
        stats = _lib.profile_stats()
        res[name] = ((b - a) * 1e3 / args.steps, {k: round(v[1] / args.steps, 4) for k, v in stats.items()})
    agree = float(np.mean(model.solid == exact))
    net_only = float(np.mean((model.logits > 0) == exact))
    at_half = float(np.mean(np.where(np.abs(model.logits) < 0.5, exact, model.logits > 0) == exact))
    ms_exact, k_exact = res["exact"]
    ms_learned, k_learned = res["learned"]
    labels_exact = sum(v for k, v in k_exact.items() if k.startswith(("qlm_kernel", "solid_bonds")))
    labels_learned = sum(v for k, v in k_learned.items() if k.startswith(("sl_features", "mlp_infer", "sl_decide")))
    print(f"[bench] exact SolidLiquid {ms_exact:.3f} ms (labels stage {labels_exact:.3f} ms); learned "
          f"{ms_learned:.3f} ms (kernels {labels_learned:.3f} ms); agreement {agree:.5f}, fallback "
          f"{model.num_fallback / n:.4f}", file=sys.stderr, flush=True)
    mlp_ms = k_learned.get("mlp_infer_kernel", 0.0)
    flops = 2.0 * n * (11 * 64 + 64 * 64 + 64)
    # dominant learned kernel: the fused feature pass, HBM-bound; compulsory
    # bytes = points 12 N + offsets 4 N + (j, d) 8 B per bond + outputs
    # (features 48 N, q_6m 104 N, q_6 4 N)
    dom = max(k_learned.items(), key=lambda kv: kv[1])[0]
    if dom.startswith("sl_features"):
        ab = 12 * n + 4 * n + 8 * len(nl) + (48 + 104 + 4) * n
        ach = ab / (k_learned[dom] * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None, "kernel": dom, "avg_ms": k_learned[dom]}
    else:
        roof = {"bound": "mfma", "achieved": None, "peak": 2500.0, "unit": "TFLOP/s", "frac": None, "traffic": None,
                "kernel": dom, "avg_ms": k_learned[dom]}
    out = {
        "metric": "pairs/s (query+RDF) & particles/s (SolidLiquid), plain vs batched/learned",
        "value": n / (ms_learned / 1e3),
        "unit": "particles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": max(args.warmup, 1),
        "ms_per_step": ms_learned,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16 (network), f32 (features, exact path)",
        "data": "synthetic (seeded two-phase FCC / liquid boxes)",
        "config": {
            "workload": "learned SolidLiquid: N=2097152 two-phase (FCC 64^3x4 at density 1.05, sigma 0.04a | "
                        "uniform liquid at 0.95); SolidLiquid l=6, q_threshold 0.7, solid_threshold 6, r_max 1.5; "
                        "MLP 11->64->64->1 BF16 MFMA, trained on 32 frames (+2 validation frames)",
            "n_points": n,
            "n_bonds": len(nl),
            "train_particles": int(model._train_size),
            "train_seconds": round(t_train, 2),
            "train_loss_first_last": [float(losses[0]), float(losses[-1])],
            "margin": model.margin,
            "agreement": agree,
            "agreement_network_only": net_only,
            "agreement_at_margin_0.5": at_half,
            "fallback_fraction": model.num_fallback / n,
            "ms_exact_solidliquid": ms_exact,
            "ms_exact_labels_kernels": labels_exact,
            "ms_learned_kernels": labels_learned,
            # like for like: the kernels that produce the solid-like labels on each side
            # (exact: q6 pass + bond dots; learned: feature pass + network + decision)
            "speedup_labels_stage": labels_exact / labels_learned if labels_learned else None,
            # not like for like: the exact side also produces ql_ij, num_connections and
            # the solid clusters (union-find), which the learned path does not
            "speedup_vs_whole_exact_solidliquid": ms_exact / ms_learned,
            "particles_per_s_exact": n / (ms_exact / 1e3),
            "kernels_exact_ms": k_exact,
            "kernels_learned_ms": k_learned,
        },
        "roofline": roof,
        "mlp_roofline": {"bound": "mfma", "achieved": round(flops / (mlp_ms * 1e-3) / 1e12, 3) if mlp_ms else None,
                         "peak": 2500.0, "unit": "TFLOP/s",
                         "frac": round(flops / (mlp_ms * 1e-3) / 1e12 / 2500.0, 5) if mlp_ms else None,
                         "kernel": "mlp_infer_kernel", "avg_ms": mlp_ms},
        "cpu_baseline": None,
    }
    return _emit(out, args)


def run_filters(args):
    """FilterSANN and FilterRAD (closed and open) over the ball NeighborList
    (r_max 1.9: ~26 candidates per point, 27M bonds) of the 1M-point frame.
    A step runs the three filters; value = candidate bonds filtered per second."""
    from freud_amd import _lib
    from freud_amd._lib import DeviceArray

    L, lib, check = _lib, _lib.lib(), _lib.check
    box, pts = make_workload()
    n = len(pts)
    d_pts = DeviceArray.from_numpy(pts)
    cb = L.make_box(box)
    nl = L.fa_nlist()
    h = ctypes.c_void_p()
    check(lib.fa_nq_create(ctypes.byref(cb), d_pts.ptr, n, L.NQ_AABB, 0.0, ctypes.byref(h)))
    try:
        ball = L.fa_query_args(L.MODE_BALL, 0xFFFFFFFF, 1.9, 0.0, -1.0, -1.0, 1)
        check(lib.fa_query(h, d_pts.ptr, n, ctypes.byref(ball), 0, ctypes.byref(nl)))
    finally:
        lib.fa_nq_destroy(h)
    nb = int(nl.num_bonds)
    outs = [L.fa_nlist() for _ in range(3)]
    d_unf = DeviceArray((n,), np.uint8)
    nu = ctypes.c_uint32()

    def step():
        check(lib.fa_filter_sann(ctypes.byref(nl), ctypes.byref(outs[0]), d_unf.ptr, ctypes.byref(nu)))
        for closed, out in ((1, outs[1]), (0, outs[2])):
            check(lib.fa_filter_rad(ctypes.byref(cb), d_pts.ptr, n, d_pts.ptr, None, ctypes.byref(nl), closed,
                                    ctypes.byref(out), d_unf.ptr, ctypes.byref(nu)))

    for w in range(max(args.warmup, 1)):
        step()
        L.synchronize()
    L.profile_reset()
    L.profile_enable(True)
    L.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    L.synchronize()
    t1 = time.perf_counter()
    L.profile_enable(False)
    stats = L.profile_stats()
    ms = (t1 - t0) * 1e3 / args.steps
    kept = [int(o.num_bonds) for o in outs]
    print(f"[bench] filters {ms:.3f} ms/step over {nb} candidate bonds: kept SANN {kept[0]}, RAD-closed {kept[1]}, "
          f"RAD-open {kept[2]}", file=sys.stderr, flush=True)
    kernels = {}
    for name, (cnt, tot) in sorted(stats.items(), key=lambda kv: -kv[1][1]):
        kernels[name] = {"launches_per_step": cnt / args.steps, "avg_ms": round(tot / cnt, 4),
                         "share": round(tot / (ms * args.steps), 4)}
    dom = max(stats.items(), key=lambda kv: kv[1][1])[0]
    cnt, tot = stats[dom]
    avg_ms = tot / cnt
    # compulsory bytes of a count pass: per candidate bond (i, j) 8 + d 4 + w 4
    # + its 2-byte code, and for RAD the point p_j 12; per query point its
    # offset 4, kept count 4, unfilled byte 1, and for RAD q_i 12
    rad = "true" in dom
    ab = nb * (8 + 4 + 4 + 2 + (12 if rad else 0)) + n * (4 + 4 + 1 + (12 if rad else 0))
    achieved = ab / (avg_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": dom, "avg_ms": round(avg_ms, 4)}
    cpu = None
    if not args.no_cpu_baseline:
        # the oracle (one core, its distance sort included) on the lists of the first 32,768 query points
        sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
        from oracle import api as oracle

        m = min(32768, n)
        end = int(L.copy_d2h(nl.offsets + 4 * m, (1,), np.uint32)[0])
        sample = {"ij": L.copy_d2h(nl.neighbors, (end, 2), np.uint32),
                  "distances": L.copy_d2h(nl.distances, (end,), np.float32),
                  "weights": L.copy_d2h(nl.weights, (end,), np.float32),
                  "vectors": L.copy_d2h(nl.vectors, (end, 3), np.float32)}
        bt = (box.Lx, box.Ly, box.Lz, box.xy, box.xz, box.yz, box.is2D)
        c0 = time.perf_counter()
        oracle.filter_sann(sample, m)
        oracle.filter_rad(bt, pts, pts[:m], sample, terminate_after_blocked=True)
        oracle.filter_rad(bt, pts, pts[:m], sample, terminate_after_blocked=False)
        c1 = time.perf_counter()
        cpu = {"value": 3 * end / (c1 - c0), "unit": "bonds/s", "cores": 1, "kind": "port",
               "sample": f"the three filters over the {end} candidate bonds of the first {m} query points"}
    out = {
        "metric": "candidate bonds filtered/s (FilterSANN + FilterRAD closed + open)",
        "value": 3 * nb / (ms / 1e3),
        "unit": "bonds/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": max(args.warmup, 1),
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 in the SANN test)",
        "data": "synthetic (seeded jittered FCC)",
        "config": {"workload": "NeighborList filters: N=1048576 jittered FCC (a=1.5874, sigma 0.05a), AABB ball "
                               "query r_max=1.9 excluding i=j; FilterSANN, FilterRAD closed, FilterRAD open",
                   "n_points": n, "n_candidate_bonds": nb, "kept": {"sann": kept[0], "rad_closed": kept[1],
                                                                     "rad_open": kept[2]},
                   "kernels": kernels},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    for o in outs + [nl]:
        lib.fa_nlist_free(ctypes.byref(o))
    return _emit(out, args)


def _emit(out, args):
    """Print a workload's JSON line, or hand it back when it runs as a
    secondary workload of the default frame run."""
    if not getattr(args, "secondary", False):
        print(json.dumps(out))
    return out


def _secondary_lines(args):
    """The batched, stress, learned and filters workloads, run after the frame
    in the default invocation (the one that also times the CPU baseline), so
    one run of bench.py measures every configuration.  Each entry is that
    workload's own line without its per-kernel table, or the error it raised."""
    import copy

    res = {}
    for name, fn in (("batched", run_batched), ("stress", run_stress), ("learned", run_learned),
                     ("filters", run_filters)):
        a = copy.copy(args)
        a.secondary = True
        a.no_cpu_baseline = True
        t0 = time.perf_counter()
        print(f"[bench] secondary workload {name}", file=sys.stderr, flush=True)
        try:
            out = fn(a)
            out = {k: v for k, v in out.items() if k != "kernels"}
        except Exception as e:  # the headline line is still printed
            out = {"error": f"{type(e).__name__}: {e}"}
        out["wall_s"] = round(time.perf_counter() - t0, 1)
        print(f"[bench] secondary workload {name} done in {out['wall_s']} s", file=sys.stderr, flush=True)
        res[name] = out
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", choices=("frame", "batched", "stress", "learned", "filters"), default="frame",
                    help="frame: one 1M-point frame (the headline line); batched: a 256-frame trajectory of uneven "
                         "frames, batched + graph-replayed against the per-frame loop; stress: 500k triclinic "
                         "(l = 4..12) + 250k 2D hexagonal (Hexatic); learned: 2M-particle two-phase SolidLiquid, "
                         "exact vs learned; filters: FilterSANN / FilterRAD over the 1M-point frame's r=1.9 list")
    ap.add_argument("--no-secondary", action="store_true",
                    help="default frame run only: skip the batched, stress, learned and filters workloads that it "
                         "otherwise runs after the frame and reports under 'secondary' (they are skipped anyway "
                         "with --no-cpu-baseline or --copy-mib)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (a known-size kernel that scales "
                         "the FETCH_SIZE / WRITE_SIZE counters of a profiled run)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    from freud_amd import _lib
    from freud_amd._lib import DeviceArray

    print(f"[bench] device: {_lib.device_name()}", file=sys.stderr, flush=True)
    if args.workload == "batched":
        return run_batched(args)
    if args.workload == "stress":
        return run_stress(args)
    if args.workload == "learned":
        return run_learned(args)
    if args.workload == "filters":
        return run_filters(args)
    if args.copy_mib > 0:
        nbytes = args.copy_mib << 20
        src = DeviceArray((nbytes // 4,), np.uint32)
        dst = DeviceArray((nbytes // 4,), np.uint32)
        src.zero()
        _lib.check(_lib.lib().fa_memcpy_d2d(dst.ptr, src.ptr, nbytes))
        _lib.synchronize()
        del src, dst
        print(f"[bench] copied {args.copy_mib} MiB device to device", file=sys.stderr, flush=True)
    box, pts = make_workload()
    n = len(pts)
    d_pts = DeviceArray.from_numpy(pts)
    fr = Frame(box, d_pts, n)
    for w in range(args.warmup):
        fr.run()
        _lib.synchronize()
        print(f"[bench] warmup {w} done", file=sys.stderr, flush=True)

    _lib.profile_reset()
    _lib.profile_enable(True)
    _lib.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        fr.run()
    _lib.synchronize()
    t1 = time.perf_counter()
    _lib.profile_enable(False)
    stats = _lib.profile_stats()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    b_ball, b_rdf, b_knn = fr.pairs()
    value = (b_ball + b_rdf) / (ms_per_step / 1e3)
    print(f"[bench] {ms_per_step:.3f} ms/step, ball bonds {b_ball}, rdf pairs {b_rdf}, knn bonds {b_knn}",
          file=sys.stderr, flush=True)

    kernels = {}
    for name, (cnt, tot) in sorted(stats.items(), key=lambda kv: -kv[1][1]):
        avg = tot / cnt
        ab = algorithmic_bytes(name, n, b_ball, b_knn)
        mb = measured_traffic(name)
        kernels[name] = {
            "launches_per_step": cnt / args.steps,
            "avg_ms": round(avg, 4),
            "share": round(tot / (ms_per_step * args.steps), 4),
            "algorithmic_bytes": ab,
            "algorithmic_frac": (None if ab is None else round(ab / (avg * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)),
            # HBM bytes from the FETCH_SIZE + WRITE_SIZE counters of a profiled run of this workload
            "measured_bytes": mb,
            "hbm_frac": (None if mb is None else round(mb / (avg * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)),
        }
    # dominant kernel by time
    dom = max(stats.items(), key=lambda kv: kv[1][1])[0]
    cnt, tot = stats[dom]
    avg_ms = tot / cnt
    if dom.startswith("(rdf") or dom.startswith("rdf"):
        roof = rdf_roofline(dom, avg_ms, rdf_stencil_evaluations([(box, n)], 5.0), b_rdf, measured_traffic(dom))
    else:
        ab = algorithmic_bytes(dom, n, b_ball, b_knn) or 0
        achieved = ab / (avg_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(dom), "kernel": dom,
                "avg_ms": round(avg_ms, 4)}

    cpu = None
    parity = None
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        cpu, ref = cpu_baseline(box, pts)
        if ref is not None:
            print("[bench] parity of the timed frame against the oracle ...", file=sys.stderr, flush=True)
            parity = frame_parity(fr, box, pts, ref)
            print(f"[bench] parity: {parity}", file=sys.stderr, flush=True)

    out = {
        "metric": "pairs/s (query+RDF) & particles/s (SolidLiquid), plain vs batched/learned",
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded FCC lattice with Gaussian jitter)",
        "config": {
            "workload": "plain single large frame: N=1048576 FCC 64^3x4, a=1.5874, sigma=0.05a; ball r_max=1.5 + "
                        "kNN k=12 + RDF 200 bins r_max=5 + Steinhardt l=6 (q6, w6, averaged q6)",
            "n_points": n,
            "ball_bonds": b_ball,
            "knn_bonds": b_knn,
            "rdf_pairs": b_rdf,
            "particles_per_s": n / (ms_per_step / 1e3),
        },
        "roofline": roof,
        "kernels": kernels,
        "cpu_baseline": cpu,
        "parity": parity,
    }
    if not (args.no_cpu_baseline or args.no_secondary or args.copy_mib > 0):
        out["secondary"] = _secondary_lines(args)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
