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
    outputs written once)."""
    nm = 2 * l + 1
    table = {
        "ball_count_kernel": 12 * n + 12 * n + 4 * n,
        "ball_fill_kernel": 12 * n + 12 * n + 4 * (n + 1) + 28 * b_ball,
        "knn_aabb_kernel": 12 * n + 12 * n + 4 * n + 8 * k * n,
        "knn_write_kernel": 8 * k * n + 4 * (n + 1) + 28 * b_knn,
        "rdf_self_kernel": 16 * n + 4 * bins,
        "qlm_kernel": 24 * b_ball + 8 * nm * n + 4 * n,
        "qlm_ave_kernel": 8 * b_ball + 8 * nm * b_ball + 8 * nm * n + 4 * n,
        "wl_kernel": 8 * nm * n + 4 * n + 4 * n,
        "grid_assign_kernel": 12 * n + 4 * n,
        "grid_scatter_kernel": 12 * n + 4 * n + 16 * n + 4 * n,
    }
    for key, v in table.items():
        if name.startswith(key) or key in name:
            return v
    return None


def cpu_baseline(box, pts, budget_s=15.0):
    """The oracle (C++ restatement of the reference's AABB tree, thread-local
    histograms and Steinhardt loops) on a bounded sample of the query points,
    on all host cores."""
    import oracle

    oracle.build()
    n = len(pts)
    cores = os.cpu_count() or 1
    bt = (box.Lx, box.Ly, box.Lz, box.xy, box.xz, box.yz, box.is2D)
    s = 2048
    elapsed = 0.0
    while True:
        q = pts[:s]
        t0 = time.perf_counter()
        nl = oracle.query(bt, pts, q, dict(r_max=1.5, exclude_ii=True), nthreads=cores)
        oracle.query(bt, pts, q, dict(num_neighbors=12, exclude_ii=True), nthreads=cores)
        counts = oracle.rdf_accumulate(bt, pts, q, 200, 5.0, exclude_ii=True, nthreads=cores)
        oracle.steinhardt(bt, pts, nl, [6], wl=True, nthreads=cores)
        oracle.steinhardt(bt, pts, nl, [6], average=True, nthreads=cores)
        elapsed = time.perf_counter() - t0
        if elapsed > budget_s / 4 or s * 2 > n:
            break
        s *= 2
    pairs = len(nl["ij"]) + int(counts.astype(np.int64).sum())
    return {
        "value": pairs / elapsed,
        "unit": "pairs/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"{s} of {n} query points of the same frame (all {n} points searched), full pipeline "
                   f"(AABB tree build + ball r=1.5 + kNN k=12 + RDF r=5 + q6/w6/avg-q6) in {elapsed:.2f} s on "
                   f"{cores} threads of {_cpu_model()}"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    from freud_amd import _lib
    from freud_amd._lib import DeviceArray

    print(f"[bench] device: {_lib.device_name()}", file=sys.stderr, flush=True)
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
        kernels[name] = {
            "launches_per_step": cnt / args.steps,
            "avg_ms": round(avg, 4),
            "share": round(tot / (ms_per_step * args.steps), 4),
            "algorithmic_bytes": ab,
            "hbm_frac": (None if ab is None else round(ab / (avg * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)),
        }
    # dominant kernel by time
    dom = max(stats.items(), key=lambda kv: kv[1][1])[0]
    cnt, tot = stats[dom]
    avg_ms = tot / cnt
    if dom.startswith("(rdf") or dom.startswith("rdf"):
        # compute-bound: ~11 flops per binned pair (3 sub, 3 mul, 2 add, sqrt, sub, mul)
        flops = 11.0 * b_rdf
        achieved = flops / (avg_ms * 1e-3) / 1e12
        roof = {"bound": "valu", "achieved": round(achieved, 3), "peak": FP32_VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / FP32_VALU_PEAK_TFLOPS, 5), "traffic": None, "kernel": dom,
                "avg_ms": round(avg_ms, 4)}
    else:
        ab = algorithmic_bytes(dom, n, b_ball, b_knn) or 0
        achieved = ab / (avg_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": dom, "avg_ms": round(avg_ms, 4)}

    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(box, pts)

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
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
